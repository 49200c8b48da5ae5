// C ABI of libmocgpu: context, grid attachment, field upload/download and the
// transport entry point.  See include/mocgpu.h for the contract.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "moc_internal.h"
#include "nlevel.h"

// Device buffer of the allocation-free entry point moc_volume_averages_noalloc:
// the volume-average partials (at most 31 x 31 ion pairs x 256 chunks x 2) and
// their chunk sums.  Each GPU the library runs on holds its own copy until the
// process ends; every call on a GPU shares it (calls that overlap on one GPU race).
constexpr size_t kScratchDoubles = (size_t)1 << 19;  // 4 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

// device address of `scratch` per device, looked up once per device (with that
// device current) by moc_create; host code uses these, never the symbol
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {};

int scratch_setup(moc_ctx* ctx) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        ctx->err = std::string("scratch buffer: device ") + std::to_string(dev) + ": " +
                   hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        double* p = nullptr;
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->err = std::string("scratch buffer: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return MOC_ERR_HIP;
        }
        g_scratchPtr[dev] = p;
    }
    ctx->d_scratch = g_scratchPtr[dev];
    return MOC_OK;
}

#define MOC_TRY(call, what)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ctx->err = std::string(what) + ": " + hipGetErrorString(e_);                 \
            return e_ == hipErrorOutOfMemory ? MOC_ERR_NOMEM : MOC_ERR_HIP;              \
        }                                                                                \
    } while (0)

template <typename P>
void dfree(P*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int dalloc(moc_ctx* ctx, void** p, size_t bytes, const char* what) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc ") + what + ": " + hipGetErrorString(e);
        *p = nullptr;
        return MOC_ERR_NOMEM;
    }
    return MOC_OK;
}

template <typename P>
int dalloc_n(moc_ctx* ctx, P** p, size_t n, const char* what) {
    return dalloc(ctx, (void**)p, n * sizeof(P), what);
}

int h2d(moc_ctx* ctx, void* d, const void* h, size_t bytes, const char* what) {
    MOC_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream), what);
    MOC_TRY(hipStreamSynchronize(ctx->stream), what);
    return MOC_OK;
}

void free_grid(MocGrid& g) {
    dfree(g.d_view64); dfree(g.d_view32); dfree(g.d_gridFirst);
    dfree(g.d_ax); dfree(g.d_ay); dfree(g.d_az);
    dfree(g.d_vx); dfree(g.d_vy); dfree(g.d_vz);
    dfree(g.d_active);
    dfree(g.d_opacity); dfree(g.d_Ne); dfree(g.d_Te); dfree(g.d_Hden);
    dfree(g.d_TeOld); dfree(g.d_NeOld);
    dfree(g.d_recPDF); dfree(g.d_totalLines); dfree(g.d_J); dfree(g.d_escaped); dfree(g.d_lineSum);
    dfree(g.d_vol);
    dfree(g.f_ax); dfree(g.f_ay); dfree(g.f_az);
    dfree(g.f_vx); dfree(g.f_vy); dfree(g.f_vz);
    dfree(g.f_opacity); dfree(g.f_Ne); dfree(g.f_recPDF); dfree(g.f_totalLines);
    dfree(g.d_ionDen);
    dfree(g.d_conv); dfree(g.d_black); dfree(g.d_nIterT); dfree(g.d_rates);
    dfree(g.d_Ndust); dfree(g.d_Tdust); dfree(g.d_absOpac); dfree(g.d_scaOpac);
    dfree(g.d_gasOpac); dfree(g.f_scaOpac); dfree(g.d_collD); dfree(g.d_peD);
    dfree(g.d_hotBin); dfree(g.d_hotCellList); dfree(g.d_hotBinList);
    dfree(g.d_hotPart); dfree(g.d_hotS);
    dfree(g.d_qT); dfree(g.d_qP); dfree(g.d_qWork);
    g = MocGrid();
}

void free_dust(MocDust& d) {
    for (void* p : d.owned) (void)hipFree(p);
    d.owned.clear();
    d = MocDust();
}

void free_balance(MocBalance& b) {
    for (void* p : b.owned) (void)hipFree(p);
    b.owned.clear();
    dfree(b.d_W);
    dfree(b.d_counters);
    b = MocBalance();
}

void free_emission(MocEmission& e) {
    for (void* p : e.owned) (void)hipFree(p);
    e.owned.clear();
    e = MocEmission();
}

// copy a host table into a new device buffer owned by the grain set
template <typename T>
int dust_put(moc_ctx* ctx, const T* h, size_t n, T** out, const char* what) {
    T* d = nullptr;
    int rc = dalloc_n(ctx, &d, n ? n : 1, what);
    if (rc) return rc;
    ctx->dust.owned.push_back((void*)d);
    if (n && h) {
        rc = h2d(ctx, d, h, n * sizeof(T), what);
        if (rc) return rc;
    }
    *out = d;
    return MOC_OK;
}

// copy a host table into a new device buffer owned by the emission set
template <typename T>
int em_put(moc_ctx* ctx, const T* h, size_t n, const T** out, const char* what) {
    T* d = nullptr;
    int rc = dalloc_n(ctx, &d, n ? n : 1, what);
    if (rc) return rc;
    ctx->emission.owned.push_back((void*)d);
    if (n) {
        if (!h) { ctx->err = std::string("moc_set_emission: missing table ") + what; return MOC_ERR_ARG; }
        rc = h2d(ctx, d, h, n * sizeof(T), what);
        if (rc) return rc;
    }
    *out = d;
    return MOC_OK;
}

// copy a host table into a new device buffer owned by the balance set
template <typename T>
int bal_put(moc_ctx* ctx, const T* h, size_t n, const T** out, const char* what) {
    T* d = nullptr;
    int rc = dalloc_n(ctx, &d, n ? n : 1, what);
    if (rc) return rc;
    ctx->balance.owned.push_back((void*)d);
    if (n) {
        if (!h) { ctx->err = std::string("moc_set_balance: missing table ") + what; return MOC_ERR_ARG; }
        rc = h2d(ctx, d, h, n * sizeof(T), what);
        if (rc) return rc;
    }
    *out = d;
    return MOC_OK;
}

// the rate buffer of the balance holds at most this many bytes (environment
// MOCGPU_RATES_BYTES overrides it): larger grids (many cells times many
// shells) run the rates and the balance in chunks
constexpr size_t kRatesBytes = (size_t)8 << 30;

size_t rates_budget() {
    const char* s = getenv("MOCGPU_RATES_BYTES");
    if (s && *s) {
        const long long v = atoll(s);
        if (v > 0) return (size_t)v;
    }
    return kRatesBytes;
}

// per-grid balance buffers (flags, iteration counts, rate matrix)
int ensure_balance_grid(moc_ctx* ctx) {
    MocGrid& g = ctx->grid;
    if (!g.attached || !ctx->balance.set) return MOC_OK;
    const size_t C1 = (size_t)g.nCells + 1;
    int rc;
    if (!g.d_conv && ((rc = dalloc_n(ctx, &g.d_conv, C1, "lgConverged")) ||
                      (rc = dalloc_n(ctx, &g.d_black, C1, "lgBlack")) ||
                      (rc = dalloc_n(ctx, &g.d_nIterT, C1, "nIterT"))))
        return rc;
    if (g.rates_cols != ctx->balance.ncol) {
        // the rates of a chunk of cells at a time, within kRatesBytes
        const size_t ncol = (size_t)ctx->balance.ncol;
        size_t chunk = rates_budget() / (8 * ncol);
        if (chunk < 64) chunk = 64;
        if (chunk > (size_t)g.nCells) chunk = g.nCells > 0 ? (size_t)g.nCells : 1;
        dfree(g.d_rates);
        if ((rc = dalloc_n(ctx, &g.d_rates, chunk * ncol, "rates"))) return rc;
        g.rates_cols = ctx->balance.ncol;
        g.rates_chunk = (int)chunk;
    }
    MOC_TRY(hipMemsetAsync(g.d_conv, 0, C1 * 4, ctx->stream), "zero flags");
    MOC_TRY(hipMemsetAsync(g.d_black, 0, C1 * 4, ctx->stream), "zero flags");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "zero flags");
    return MOC_OK;
}

void free_atomic(MocAtomic& a) {
    dfree(a.d_elemZ); dfree(a.d_specElem); dfree(a.d_specIon);
    dfree(a.d_abund); dfree(a.d_sigma); dfree(a.d_brems); dfree(a.d_log10nu);
    a = MocAtomic();
}

int64_t ionden_count(moc_ctx* ctx) {
    return ((int64_t)ctx->grid.nCells + 1) * ctx->atomic.nElemUsed * ctx->atomic.nstages;
}

// the ionDen plane set exists once both the grid and the element set are known
int ensure_ionden(moc_ctx* ctx) {
    MocGrid& g = ctx->grid;
    if (!g.attached || !ctx->atomic.set || g.d_ionDen) return MOC_OK;
    const size_t n = (size_t)ionden_count(ctx);
    int rc = dalloc_n(ctx, &g.d_ionDen, n, "ionDen");
    if (rc) return rc;
    MOC_TRY(hipMemsetAsync(g.d_ionDen, 0, n * 8, ctx->stream), "zero ionDen");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "zero ionDen");
    return MOC_OK;
}

// cell widths of the volume element (photon_mod.f90:2169-2223), in 1e15 cm:
// interior cells span the two half-gaps; the first cell is half-width in a
// symmetric (octant) or plane-parallel grid, the last cell in a plane-parallel
// grid only (full-width otherwise), as the reference has it.
std::vector<double> widths(const std::vector<double>& a, bool sym, bool plane) {
    const int n = (int)a.size();
    std::vector<double> w(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double d;
        if (n == 1) d = 0.0;
        else if (i > 0 && i < n - 1) d = fabs(a[i + 1] - a[i - 1]) / 2.0;
        else if (i == 0) d = (sym || plane) ? fabs(a[1] - a[0]) / 2.0 : fabs(a[1] - a[0]);
        else d = plane ? fabs(a[i] - a[i - 1]) / 2.0 : fabs(a[i] - a[i - 1]);
        w[i] = d / 1.e15;
    }
    return w;
}

bool is_cellbin(int f) {
    return f == MOC_F_OPACITY || f == MOC_F_JSTE || f == MOC_F_RECPDF || f == MOC_F_ESCAPED ||
           f == MOC_F_DUSTPDF || f == MOC_F_ABSOPAC || f == MOC_F_SCAOPAC;
}

double* field_ptr(moc_ctx* ctx, int f) {
    MocGrid& g = ctx->grid;
    switch (f) {
        case MOC_F_OPACITY: return g.d_opacity;
        case MOC_F_JSTE: return g.d_J;
        case MOC_F_RECPDF: return g.d_recPDF;
        case MOC_F_TOTALLINES: return g.d_totalLines;
        case MOC_F_NE: return g.d_Ne;
        case MOC_F_TE: return g.d_Te;
        case MOC_F_HDEN: return g.d_Hden;
        case MOC_F_ESCAPED: return g.d_escaped;
        case MOC_F_IONDEN: return g.d_ionDen;
        case MOC_F_NDUST: return g.d_Ndust;
        case MOC_F_TDUST: return g.d_Tdust;
        case MOC_F_DUSTPDF: return ctx->dust.set && !ctx->dust.lgGas ? g.d_recPDF : nullptr;
        case MOC_F_ABSOPAC: return g.d_absOpac;
        case MOC_F_SCAOPAC: return g.d_scaOpac;
        case MOC_F_GRAINCOLL: return g.d_collD;
        case MOC_F_GRAINPE: return g.d_peD;
        default: return nullptr;
    }
}

// device storage order of a (cell, bin) field: recPDF is bin-fastest (rows
// searched per cell), the others cell-fastest (planes walked per bin)
int device_layout(int f) {
    return f == MOC_F_RECPDF || f == MOC_F_DUSTPDF ? MOC_LAYOUT_CELLROW : MOC_LAYOUT_FORTRAN;
}

int64_t field_count(moc_ctx* ctx, int f) {
    const int64_t C1 = ctx->grid.nCells + 1;
    if (is_cellbin(f)) return C1 * ctx->nbins;
    if (f == MOC_F_IONDEN) return ionden_count(ctx);
    if (f == MOC_F_TDUST) return C1 * (ctx->dust.nA + 1) * (ctx->dust.nS + 1);
    if (f == MOC_F_GRAINCOLL || f == MOC_F_GRAINPE) return C1 * ctx->dust.nA * ctx->dust.nS;
    return C1;
}

int refresh_f32(moc_ctx* ctx) {
    MocGrid& g = ctx->grid;
    if (!g.f32_dirty) return MOC_OK;
    const int64_t C1 = g.nCells + 1, B = ctx->nbins;
    if (!g.op32_master) moc_convert_f64_to_f32(g.d_opacity, g.f_opacity, C1 * B, ctx->stream);
    if (!g.pdf32_master) moc_convert_f64_to_f32(g.d_recPDF, g.f_recPDF, C1 * B, ctx->stream);
    moc_convert_f64_to_f32(g.d_Ne, g.f_Ne, C1, ctx->stream);
    moc_convert_f64_to_f32(g.d_totalLines, g.f_totalLines, C1, ctx->stream);
    if (g.d_scaOpac) moc_convert_f64_to_f32(g.d_scaOpac, g.f_scaOpac, C1 * B, ctx->stream);
    MOC_TRY(hipGetLastError(), "f32 refresh");
    g.f32_dirty = false;
    return MOC_OK;
}

// Bring the FP64 plane of field f up to date where the mixed form's FP32
// kernels wrote its FP32 plane (the widening is exact, so both then agree and
// the FP64 plane is the master again).
int sync_f64_master(moc_ctx* ctx, int f) {
    MocGrid& g = ctx->grid;
    const int64_t n = ((int64_t)g.nCells + 1) * ctx->nbins;
    if (f == MOC_F_OPACITY && g.op32_master) {
        moc_convert_f32_to_f64(g.f_opacity, g.d_opacity, n, ctx->stream);
        g.op32_master = false;
    }
    if (f == MOC_F_RECPDF && g.pdf32_master) {
        moc_convert_f32_to_f64(g.f_recPDF, g.d_recPDF, n, ctx->stream);
        g.pdf32_master = false;
    }
    MOC_TRY(hipGetLastError(), "FP64 plane refresh");
    return MOC_OK;
}

int attach_grids(moc_ctx* ctx, int nG, const moc_grid_desc* gds, const char* who, bool nested = false);
int upload_nested_views(moc_ctx* ctx, int s);
MocBalanceArgs balance_args(moc_ctx* ctx, const double** W);
void drop_graph(moc_ctx* ctx);  // a captured iteration holds the arguments of its time
// nested sub-grids: the view of every grid for source s (0-based): the
// source's deltaE, the mother's symmetry and R_out in every view, the start
// cell in the source's grid
int upload_nested_views(moc_ctx* ctx, int s) {
    MocGrid& g = ctx->grid;
    const int nG = g.nGrids;
    const double dE = ctx->Lstar[s] / ctx->nPhotons[s];  // photon_mod.f90:88
    if (!(dE >= 0.0)) { ctx->err = "moc_transport: deltaEUsed < 0"; return MOC_ERR_ARG; }
    std::vector<MocGridView<double>> v64((size_t)nG);
    std::vector<MocGridView<float>> v32((size_t)nG);
    for (int k = 0; k < nG; ++k) {
        MocGridView<double>& a = v64[k];
        MocGridView<float>& b = v32[k];
        a.ax = g.d_ax + g.goffx[k]; a.ay = g.d_ay + g.goffy[k]; a.az = g.d_az + g.goffz[k];
        a.vx = g.d_vx + g.goffx[k]; a.vy = g.d_vy + g.goffy[k]; a.vz = g.d_vz + g.goffz[k];
        b.ax = g.f_ax + g.goffx[k]; b.ay = g.f_ay + g.goffy[k]; b.az = g.f_az + g.goffz[k];
        b.vx = g.f_vx + g.goffx[k]; b.vy = g.f_vy + g.goffy[k]; b.vz = g.f_vz + g.goffz[k];
        a.active = b.active = (k > 0 && g.gdense[k]) ? nullptr : g.d_active + g.goffa[k];
        a.sourceCDF = ctx->d_sourceCDF + (size_t)s * ctx->nbins;
        b.sourceCDF = ctx->f_sourceCDF + (size_t)s * ctx->nbins;
        a.nx = b.nx = g.gnx[k]; a.ny = b.ny = g.gny[k]; a.nz = b.nz = g.gnz[k];
        a.base = b.base = g.gbase[k];
        a.sym = b.sym = g.sym;
        a.plane = b.plane = 0;
        a.sub = b.sub = k > 0 ? 1 : 0;
        for (int d = 0; d < 3; ++d) {
            a.starIdx[d] = b.starIdx[d] = g.nestStartIdx[d];
            a.star[d] = ctx->starPos[3 * s + d];
            b.star[d] = (float)ctx->starPos[3 * s + d];
        }
        a.R_out = g.gR_out[0]; b.R_out = (float)g.gR_out[0];
        a.deltaE = dE; b.deltaE = (float)dE;
    }
    int rc;
    if ((rc = h2d(ctx, g.d_view64, v64.data(), v64.size() * sizeof(MocGridView<double>), "grid views")) ||
        (rc = h2d(ctx, g.d_view32, v32.data(), v32.size() * sizeof(MocGridView<float>), "grid views")))
        return rc;
    return MOC_OK;
}

int run_transport(moc_ctx* ctx, int iStar, int64_t n, uint64_t seed, uint32_t ctr_base,
                  moc_transport_stats* stats);

}  // namespace

extern "C" {

const char* moc_version(void) { return "gfx950 mocgpu 0.1"; }

const char* moc_last_error(const moc_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int moc_create(moc_ctx** out, const moc_tables* t) {
    if (!out || !t) return MOC_ERR_ARG;
    *out = nullptr;
    moc_ctx* ctx = new moc_ctx();
    *out = ctx;
    const int B = t->nbins;
    if (B < 2 || !t->nuArray || !t->widFlx || t->nStars < 1 || !t->sourceCDF || !t->Lstar ||
        !t->nPhotons || !t->starPosition || !t->starIndex || t->nAngles < 2 || !t->KNsigmaT ||
        !t->KNsigmaCDF || !t->Pcomp ||
        (t->precision != MOC_PREC_FP64 && t->precision != MOC_PREC_MIXED)) {
        ctx->err = "moc_create: invalid table set";
        return MOC_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        ctx->err = "moc_create: no HIP device";
        return MOC_ERR_HIP;
    }
    ctx->precision = t->precision;
    ctx->nbins = B;
    ctx->nAngles = t->nAngles;
    ctx->lgCompton = t->lgCompton;
    ctx->nStars = t->nStars;
    ctx->ionEdge1 = t->ionEdge1;
    ctx->nuArray.assign(t->nuArray, t->nuArray + B);
    ctx->widFlx.assign(t->widFlx, t->widFlx + B);
    ctx->Lstar.assign(t->Lstar, t->Lstar + t->nStars);
    ctx->nPhotons.assign(t->nPhotons, t->nPhotons + t->nStars);
    ctx->starPos.assign(t->starPosition, t->starPosition + 3 * t->nStars);
    ctx->starIdx.assign(t->starIndex, t->starIndex + 3 * t->nStars);
    for (int i = 1; i < B; ++i)
        if (!(ctx->nuArray[i] > ctx->nuArray[i - 1])) {
            ctx->err = "moc_create: nuArray must be strictly increasing";
            return MOC_ERR_ARG;
        }
    MOC_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream");
    int rc0 = scratch_setup(ctx);
    if (rc0) return rc0;
    MOC_TRY(hipEventCreate(&ctx->ev0), "event");
    MOC_TRY(hipEventCreate(&ctx->ev1), "event");
    const size_t nA = (size_t)B * (size_t)t->nAngles, nS = (size_t)B * (size_t)t->nStars;
    int rc;
    if ((rc = dalloc_n(ctx, &ctx->d_nu, B, "nu")) || (rc = dalloc_n(ctx, &ctx->d_KNsigmaT, B, "KN")) ||
        (rc = dalloc_n(ctx, &ctx->d_KNcdf, nA, "KNcdf")) || (rc = dalloc_n(ctx, &ctx->d_Pcomp, nA, "Pcomp")) ||
        (rc = dalloc_n(ctx, &ctx->d_sourceCDF, nS, "srcCDF")) ||
        (rc = dalloc_n(ctx, &ctx->f_nu, B, "nu32")) || (rc = dalloc_n(ctx, &ctx->f_KNsigmaT, B, "KN32")) ||
        (rc = dalloc_n(ctx, &ctx->f_KNcdf, nA, "KNcdf32")) || (rc = dalloc_n(ctx, &ctx->f_Pcomp, nA, "Pcomp32")) ||
        (rc = dalloc_n(ctx, &ctx->f_sourceCDF, nS, "srcCDF32")) ||
        (rc = dalloc_n(ctx, &ctx->d_Qphot, 1, "Qphot")) ||
        (rc = dalloc_n(ctx, &ctx->d_counters, MOC_CNT_N, "counters")))
        return rc;
    if ((rc = h2d(ctx, ctx->d_nu, t->nuArray, B * 8, "nu")) ||
        (rc = h2d(ctx, ctx->d_KNsigmaT, t->KNsigmaT, B * 8, "KN")) ||
        (rc = h2d(ctx, ctx->d_KNcdf, t->KNsigmaCDF, nA * 8, "KNcdf")) ||
        (rc = h2d(ctx, ctx->d_Pcomp, t->Pcomp, nA * 8, "Pcomp")) ||
        (rc = h2d(ctx, ctx->d_sourceCDF, t->sourceCDF, nS * 8, "srcCDF")))
        return rc;
    moc_convert_f64_to_f32(ctx->d_nu, ctx->f_nu, B, ctx->stream);
    moc_convert_f64_to_f32(ctx->d_KNsigmaT, ctx->f_KNsigmaT, B, ctx->stream);
    moc_convert_f64_to_f32(ctx->d_KNcdf, ctx->f_KNcdf, (int64_t)nA, ctx->stream);
    moc_convert_f64_to_f32(ctx->d_Pcomp, ctx->f_Pcomp, (int64_t)nA, ctx->stream);
    moc_convert_f64_to_f32(ctx->d_sourceCDF, ctx->f_sourceCDF, (int64_t)nS, ctx->stream);
    MOC_TRY(hipStreamSynchronize(ctx->stream), "table conversion");
    return MOC_OK;
}

void moc_destroy(moc_ctx* ctx) {
    if (!ctx) return;
    drop_graph(ctx);
    dfree(ctx->d_ctr);
    free_grid(ctx->grid);
    free_atomic(ctx->atomic);
    free_balance(ctx->balance);
    free_emission(ctx->emission);
    free_dust(ctx->dust);
    dfree(ctx->d_nu); dfree(ctx->d_KNsigmaT); dfree(ctx->d_KNcdf); dfree(ctx->d_Pcomp);
    dfree(ctx->d_sourceCDF);
    dfree(ctx->f_nu); dfree(ctx->f_KNsigmaT); dfree(ctx->f_KNcdf); dfree(ctx->f_Pcomp);
    dfree(ctx->f_sourceCDF);
    dfree(ctx->d_Qphot); dfree(ctx->d_counters);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int moc_grid_attach(moc_ctx* ctx, int ig, const moc_grid_desc* gd) {
    if (!ctx) return MOC_ERR_ARG;
    if (ig != 1) {
        ctx->err = "moc_grid_attach: invalid grid (one grid, ig=1, n>=2 per axis)";
        return MOC_ERR_ARG;
    }
    return attach_grids(ctx, 1, gd, "moc_grid_attach");
}

int moc_grid_attach_batch(moc_ctx* ctx, int nGrids, const moc_grid_desc* grids) {
    if (!ctx) return MOC_ERR_ARG;
    if (nGrids < 1 || !grids) { ctx->err = "moc_grid_attach_batch: nGrids >= 1 grids"; return MOC_ERR_ARG; }
    return attach_grids(ctx, nGrids, grids, "moc_grid_attach_batch");
}

int moc_grid_attach_nested(moc_ctx* ctx, int nGrids, const moc_grid_desc* grids) {
    if (!ctx) return MOC_ERR_ARG;
    if (nGrids < 2 || !grids) {
        ctx->err = "moc_grid_attach_nested: a mother grid and at least one sub-grid";
        return MOC_ERR_ARG;
    }
    return attach_grids(ctx, nGrids, grids, "moc_grid_attach_nested", true);
}

}  // extern "C"

namespace {

// Attach nG grids; more than one is a batch of independent grids whose cells
// are numbered one grid after the other and whose axes and active maps are
// concatenated (active ids become global).
int attach_grids(moc_ctx* ctx, int nG, const moc_grid_desc* gds, const char* who, bool nested) {
    drop_graph(ctx);
    const std::string w(who);
    for (int k = 0; k < nG; ++k) {
        const moc_grid_desc* gd = gds + k;
        if (gd->nx < 2 || gd->ny < 2 || gd->nz < 2 || !gd->xAxis || !gd->yAxis || !gd->zAxis ||
            !gd->active || gd->nCells < 1) {
            ctx->err = w + ": invalid grid " + std::to_string(k) + " (n>=2 per axis, nCells>=1)";
            return MOC_ERR_ARG;
        }
        if (nested && (gds->lgPlaneIonization || gd->lgPlaneIonization)) {
            ctx->err = w + ": nested sub-grids of a plane-parallel slab are not supported";
            return MOC_ERR_ARG;
        }
        if (nested && k > 0) {
            // a sub-grid lies inside its mother's box (setSubGrids, grid_mod.f90:1834-2100)
            const moc_grid_desc* m = gds;
            if (gd->xAxis[0] < m->xAxis[0] || gd->xAxis[gd->nx - 1] > m->xAxis[m->nx - 1] ||
                gd->yAxis[0] < m->yAxis[0] || gd->yAxis[gd->ny - 1] > m->yAxis[m->ny - 1] ||
                gd->zAxis[0] < m->zAxis[0] || gd->zAxis[gd->nz - 1] > m->zAxis[m->nz - 1]) {
                ctx->err = w + ": sub-grid " + std::to_string(k + 1) + " reaches outside the mother grid";
                return MOC_ERR_ARG;
            }
            continue;  // a sub-grid takes the mother's symmetry
        }
        if ((gd->lgSymmetricXYZ ? 1 : 0) != (gds->lgSymmetricXYZ ? 1 : 0) ||
            (gd->lgPlaneIonization ? 1 : 0) != (gds->lgPlaneIonization ? 1 : 0)) {
            ctx->err = w + ": the grids of a batch share lgSymmetricXYZ and lgPlaneIonization";
            return MOC_ERR_ARG;
        }
        // set_input_mod.f90:581-584 and :622-625 (the reference drops the
        // symmetry with a warning; here the caller chooses)
        if (gd->lgPlaneIonization && gd->lgSymmetricXYZ) {
            ctx->err = w + ": lgPlaneIonization and lgSymmetricXYZ are exclusive";
            return MOC_ERR_ARG;
        }
        if (gd->lgPlaneIonization && (gd->nx < 3 || gd->nz < 3)) {
            ctx->err = w + ": a plane-parallel grid is at least 3 cells deep in x and z";
            return MOC_ERR_ARG;
        }
    }
    if (nested && ctx->nStars != 1) {
        ctx->err = w + ": nested sub-grids take a single source";
        return MOC_ERR_ARG;
    }
    if (nG > 1 && !nested && ctx->nStars != nG) {
        ctx->err = w + ": a batch of " + std::to_string(nG) + " grids needs one star per grid (" +
                   std::to_string(ctx->nStars) + " stars)";
        return MOC_ERR_ARG;
    }
    int64_t total = 0;
    for (int k = 0; k < nG; ++k) total += gds[k].nCells;
    if (total > 2000000000LL) { ctx->err = w + ": more than 2e9 cells"; return MOC_ERR_ARG; }
    free_grid(ctx->grid);
    free_dust(ctx->dust);  // its per-cell arrays belonged to the old grid
    MocGrid& g = ctx->grid;
    g.nGrids = nG;
    g.nested = nested ? 1 : 0;
    g.nx = gds->nx; g.ny = gds->ny; g.nz = gds->nz;
    g.nCells = (int)total;
    g.sym = gds->lgSymmetricXYZ ? 1 : 0;
    g.plane = gds->lgPlaneIonization ? 1 : 0;
    g.R_out = gds->R_out;
    std::vector<double> vx, vy, vz;
    std::vector<double> vol((size_t)total + 1, 0.0);  // getVolume of every cell
    int base = 0;
    for (int k = 0; k < nG; ++k) {
        const moc_grid_desc* gd = gds + k;
        const int nx = gd->nx, ny = gd->ny, nz = gd->nz;
        g.gnx.push_back(nx); g.gny.push_back(ny); g.gnz.push_back(nz);
        g.gbase.push_back(base);
        g.gR_out.push_back(gd->R_out);
        g.goffx.push_back(g.ax.size()); g.goffy.push_back(g.ay.size());
        g.goffz.push_back(g.az.size()); g.goffa.push_back(g.active.size());
        const std::vector<double> ax(gd->xAxis, gd->xAxis + nx), ay(gd->yAxis, gd->yAxis + ny),
            az(gd->zAxis, gd->zAxis + nz);
        g.ax.insert(g.ax.end(), ax.begin(), ax.end());
        g.ay.insert(g.ay.end(), ay.begin(), ay.end());
        g.az.insert(g.az.end(), az.begin(), az.end());
        // a sub-grid's end cells are half-width on every face (photon_mod.f90:2169-2223, gP > 1)
        const bool half = nested && k > 0;
        const std::vector<double> wx = widths(ax, g.sym || half, g.plane || half),
                                  wy = widths(ay, g.sym || half, g.plane || half),
                                  wz = widths(az, g.sym || half, g.plane || half);
        vx.insert(vx.end(), wx.begin(), wx.end());
        vy.insert(vy.end(), wy.begin(), wy.end());
        vz.insert(vz.end(), wz.begin(), wz.end());
        // every cell active and numbered with z fastest, then y, then x: the id
        // is computed instead of looked up in the transport
        const size_t n3 = (size_t)nx * ny * nz;
        bool dense = true;
        for (size_t i = 0; i < n3; ++i) {
            const int id = gd->active[i];
            // the mother of nested grids marks the cells of sub-grid g as -g (fillGrid, grid_mod.f90:907-960)
            const bool subCell = nested && k == 0 && id <= -2 && id >= -nG;
            if ((id < 0 && !subCell) || id > gd->nCells) {
                ctx->err = w + (nested && k == 0 ? ": mother active ids must lie in [0, nCells] or be -g for sub-grid g in 2..nGrids"
                                                 : ": active ids must lie in [0, nCells] (sub-grids: moc_grid_attach_nested)");
                free_grid(g);
                return MOC_ERR_ARG;
            }
            const size_t ix = i % nx, iy = (i / nx) % ny, iz = i / ((size_t)nx * ny);
            if (id != 1 + (int)(iz + nz * (iy + ny * ix))) dense = false;
            g.active.push_back(id > 0 ? base + id : (subCell ? id : 0));
            if (id > 0) vol[(size_t)base + id] = wx[ix] * wy[iy] * wz[iz];
        }
        g.gdense.push_back(dense ? 1 : 0);
        if (nG > 1 && !nested) {  // star k lights grid k
            const int* si = &ctx->starIdx[3 * k];
            if (si[0] < 1 || si[0] > nx || si[1] < 1 || si[1] > ny || si[2] < 1 || si[2] > nz) {
                ctx->err = w + ": star cell index outside its grid " + std::to_string(k);
                free_grid(g);
                return MOC_ERR_ARG;
            }
        }
        base += gd->nCells;
    }
    g.dense = nG == 1 && g.gdense[0];
    if (nested) {
        // the source's grid: the mother, or the sub-grid that covers its cell
        // (setStarPosition, grid_mod.f90:3691-3788)
        const int* si = ctx->starIdx.data();
        if (si[0] < 1 || si[0] > g.nx || si[1] < 1 || si[1] > g.ny || si[2] < 1 || si[2] > g.nz) {
            ctx->err = w + ": star cell index outside the mother grid";
            free_grid(g);
            return MOC_ERR_ARG;
        }
        const int id = g.active[(size_t)(si[0] - 1) + (size_t)g.nx * ((size_t)(si[1] - 1) + (size_t)g.ny * (si[2] - 1))];
        g.nestStart = 0;
        for (int d = 0; d < 3; ++d) g.nestStartIdx[d] = si[d] - 1;
        if (id < 0) {
            const int k = -id - 1;
            g.nestStart = k;
            const double* axes[3] = {g.ax.data() + g.goffx[k], g.ay.data() + g.goffy[k], g.az.data() + g.goffz[k]};
            const int nn[3] = {g.gnx[k], g.gny[k], g.gnz[k]};
            for (int d = 0; d < 3; ++d) {  // locate, then the next cell above the midpoint
                const double x = ctx->starPos[d];
                const double* a = axes[d];
                int p = 0;
                while (p + 1 < nn[d] && a[p + 1] <= x) ++p;
                if (p < nn[d] - 1 && x > (a[p] + a[p + 1]) / 2.0) ++p;
                g.nestStartIdx[d] = p;
            }
        }
    }
    if (nG == 1)
        for (int s = 0; s < ctx->nStars; ++s) {
            const int* si = &ctx->starIdx[3 * s];
            if (si[0] < 1 || si[0] > g.nx || si[1] < 1 || si[1] > g.ny || si[2] < 1 || si[2] > g.nz) {
                ctx->err = w + ": star cell index outside the grid";
                free_grid(g);
                return MOC_ERR_ARG;
            }
        }
    const size_t C1 = (size_t)g.nCells + 1, B = (size_t)ctx->nbins;
    const size_t NX = g.ax.size(), NY = g.ay.size(), NZ = g.az.size(), NA = g.active.size();
    int rc;
    if ((rc = dalloc_n(ctx, &g.d_ax, NX, "ax")) || (rc = dalloc_n(ctx, &g.d_ay, NY, "ay")) ||
        (rc = dalloc_n(ctx, &g.d_az, NZ, "az")) || (rc = dalloc_n(ctx, &g.d_vx, NX, "vx")) ||
        (rc = dalloc_n(ctx, &g.d_vy, NY, "vy")) || (rc = dalloc_n(ctx, &g.d_vz, NZ, "vz")) ||
        (rc = dalloc_n(ctx, &g.d_active, NA, "active")) ||
        (rc = dalloc_n(ctx, &g.d_opacity, C1 * B, "opacity")) ||
        (rc = dalloc_n(ctx, &g.d_Ne, C1, "Ne")) || (rc = dalloc_n(ctx, &g.d_Te, C1, "Te")) ||
        (rc = dalloc_n(ctx, &g.d_Hden, C1, "Hden")) ||
        (rc = dalloc_n(ctx, &g.d_TeOld, C1, "TeOld")) || (rc = dalloc_n(ctx, &g.d_NeOld, C1, "NeOld")) ||
        (rc = dalloc_n(ctx, &g.d_recPDF, C1 * B, "recPDF")) ||
        (rc = dalloc_n(ctx, &g.d_totalLines, C1, "totalLines")) ||
        (rc = dalloc_n(ctx, &g.d_lineSum, C1, "line sums")) ||
        (rc = dalloc_n(ctx, &g.d_vol, C1, "cell volumes")) ||
        (rc = dalloc_n(ctx, &g.d_J, C1 * B, "J")) ||
        (rc = dalloc_n(ctx, &g.d_escaped, C1 * B, "escaped")) ||
        (rc = dalloc_n(ctx, &g.f_ax, NX, "ax32")) || (rc = dalloc_n(ctx, &g.f_ay, NY, "ay32")) ||
        (rc = dalloc_n(ctx, &g.f_az, NZ, "az32")) || (rc = dalloc_n(ctx, &g.f_vx, NX, "vx32")) ||
        (rc = dalloc_n(ctx, &g.f_vy, NY, "vy32")) || (rc = dalloc_n(ctx, &g.f_vz, NZ, "vz32")) ||
        (rc = dalloc_n(ctx, &g.f_opacity, C1 * B, "opacity32")) ||
        (rc = dalloc_n(ctx, &g.f_Ne, C1, "Ne32")) ||
        (rc = dalloc_n(ctx, &g.f_recPDF, C1 * B, "recPDF32")) ||
        (rc = dalloc_n(ctx, &g.f_totalLines, C1, "totalLines32")) ||
        (nG > 1 && ((rc = dalloc_n(ctx, &g.d_view64, (size_t)nG, "grid views")) ||
                    (rc = dalloc_n(ctx, &g.d_view32, (size_t)nG, "grid views")) ||
                    (rc = dalloc_n(ctx, &g.d_gridFirst, (size_t)nG + 1, "grid packets"))))) {
        free_grid(g);
        return rc;
    }
    if ((rc = h2d(ctx, g.d_ax, g.ax.data(), NX * 8, "ax")) ||
        (rc = h2d(ctx, g.d_ay, g.ay.data(), NY * 8, "ay")) ||
        (rc = h2d(ctx, g.d_az, g.az.data(), NZ * 8, "az")) ||
        (rc = h2d(ctx, g.d_vx, vx.data(), NX * 8, "vx")) ||
        (rc = h2d(ctx, g.d_vy, vy.data(), NY * 8, "vy")) ||
        (rc = h2d(ctx, g.d_vz, vz.data(), NZ * 8, "vz")) ||
        (rc = h2d(ctx, g.d_active, g.active.data(), NA * 4, "active")) ||
        (rc = h2d(ctx, g.d_vol, vol.data(), C1 * 8, "cell volumes"))) {
        free_grid(g);
        return rc;
    }
    MOC_TRY(hipMemsetAsync(g.d_opacity, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_recPDF, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_J, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_escaped, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_Ne, 0, C1 * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_Te, 0, C1 * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_Hden, 0, C1 * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_totalLines, 0, C1 * 8, ctx->stream), "zero");
    moc_convert_f64_to_f32(g.d_ax, g.f_ax, (int64_t)NX, ctx->stream);
    moc_convert_f64_to_f32(g.d_ay, g.f_ay, (int64_t)NY, ctx->stream);
    moc_convert_f64_to_f32(g.d_az, g.f_az, (int64_t)NZ, ctx->stream);
    moc_convert_f64_to_f32(g.d_vx, g.f_vx, (int64_t)NX, ctx->stream);
    moc_convert_f64_to_f32(g.d_vy, g.f_vy, (int64_t)NY, ctx->stream);
    moc_convert_f64_to_f32(g.d_vz, g.f_vz, (int64_t)NZ, ctx->stream);
    MOC_TRY(hipStreamSynchronize(ctx->stream), "grid attach");
    g.f32_dirty = true;
    g.attached = true;
    int rc2 = ensure_ionden(ctx);
    if (rc2) return rc2;
    return ensure_balance_grid(ctx);
}

}  // namespace

extern "C" {

int moc_set_atomic(moc_ctx* ctx, const moc_atomic* at) {
    if (!ctx) return MOC_ERR_ARG;
    drop_graph(ctx);
    if (!at || at->nElemUsed < 1 || at->nElemUsed > 30 || at->nstages < 1 || at->nstages > 31 ||
        at->nSpecies < 0 || !at->elemZ || !at->abund || (at->nSpecies > 0 && (!at->speciesElem ||
        !at->speciesIon || !at->sigma)) || !at->brems) {
        ctx->err = "moc_set_atomic: invalid element set";
        return MOC_ERR_ARG;
    }
    for (int s = 0; s < at->nSpecies; ++s)
        if (at->speciesElem[s] < 0 || at->speciesElem[s] >= at->nElemUsed || at->speciesIon[s] < 1 ||
            at->speciesIon[s] > at->nstages) {
            ctx->err = "moc_set_atomic: species index outside the element/stage range";
            return MOC_ERR_ARG;
        }
    const bool same_shape = ctx->atomic.set && ctx->atomic.nElemUsed == at->nElemUsed &&
                            ctx->atomic.nstages == at->nstages;
    free_atomic(ctx->atomic);
    // the balance and emission sets point into the element set: both must be set again
    free_balance(ctx->balance);
    free_emission(ctx->emission);
    if (!same_shape) dfree(ctx->grid.d_ionDen);
    MocAtomic& a = ctx->atomic;
    a.nElemUsed = at->nElemUsed;
    a.nstages = at->nstages;
    a.nSpecies = at->nSpecies;
    a.elemZ.assign(at->elemZ, at->elemZ + a.nElemUsed);
    a.abund.assign(at->abund, at->abund + a.nElemUsed);
    a.specElem.assign(at->speciesElem, at->speciesElem + a.nSpecies);
    a.specIon.assign(at->speciesIon, at->speciesIon + a.nSpecies);
    const size_t B = ctx->nbins, S = a.nSpecies;
    std::vector<double> lnu(B);
    for (size_t b = 0; b < B; ++b) lnu[b] = log10(ctx->nuArray[b]);
    int rc;
    if ((rc = dalloc_n(ctx, &a.d_elemZ, a.nElemUsed, "elemZ")) ||
        (rc = dalloc_n(ctx, &a.d_abund, a.nElemUsed, "abund")) ||
        (rc = dalloc_n(ctx, &a.d_specElem, S, "specElem")) ||
        (rc = dalloc_n(ctx, &a.d_specIon, S, "specIon")) ||
        (rc = dalloc_n(ctx, &a.d_sigma, S * B, "sigma")) ||
        (rc = dalloc_n(ctx, &a.d_brems, B, "brems")) ||
        (rc = dalloc_n(ctx, &a.d_log10nu, B, "log10nu")))
        return rc;
    if ((rc = h2d(ctx, a.d_elemZ, at->elemZ, a.nElemUsed * 4, "elemZ")) ||
        (rc = h2d(ctx, a.d_abund, at->abund, a.nElemUsed * 8, "abund")) ||
        (S && (rc = h2d(ctx, a.d_specElem, at->speciesElem, S * 4, "specElem"))) ||
        (S && (rc = h2d(ctx, a.d_specIon, at->speciesIon, S * 4, "specIon"))) ||
        (S && (rc = h2d(ctx, a.d_sigma, at->sigma, S * B * 8, "sigma"))) ||
        (rc = h2d(ctx, a.d_brems, at->brems, B * 8, "brems")) ||
        (rc = h2d(ctx, a.d_log10nu, lnu.data(), B * 8, "log10nu")))
        return rc;
    a.set = true;
    return ensure_ionden(ctx);
}

int run_opacity(moc_ctx* ctx, bool skipUnchanged, float* ms);
int run_emission(moc_ctx* ctx, bool skipUnchanged, float* ms);

MocDustArgs dust_args(moc_ctx* ctx) {
    MocGrid& g = ctx->grid;
    MocDust& d = ctx->dust;
    MocDustArgs a{};
    a.nCells = g.nCells; a.B = ctx->nbins; a.nS = d.nS; a.nA = d.nA; a.nT = d.nT;
    a.dustOnly = d.lgGas ? 0 : 1;
    a.abun = d.d_abun; a.weight = d.d_weight; a.tsub = d.d_tsub;
    a.Cabs = d.d_Cabs; a.Csca = d.d_Csca; a.emInt = d.d_emInt;
    a.nu = ctx->d_nu; a.widFlx = d.d_widFlx;
    a.Ndust = g.d_Ndust; a.Tdust = g.d_Tdust; a.absOpac = g.d_absOpac; a.scaOpac = g.d_scaOpac;
    a.gasOpac = d.lgGas ? g.d_gasOpac : nullptr;
    a.opacity = g.d_opacity;
    a.pdf = g.d_recPDF;
    a.R = d.d_R;
    a.collD = d.lgPE ? g.d_collD : nullptr;
    a.peD = d.lgPE ? g.d_peD : nullptr;
    a.lgConverged = g.d_conv; a.lgBlack = g.d_black;
    a.counters = d.d_counters;
    a.XHILimit = d.XHILimit;
    a.qActive = d.lgQHeat && d.qActive && d.nQ > 0 ? 1 : 0;
    a.nQ = d.nQ; a.nTb = d.nTb; a.qIdx = d.d_qIdx; a.qG = d.d_qG; a.kind = d.d_kind;
    a.radius = d.d_radius; a.Tmax = d.Tmax; a.J = g.d_J;
    a.qT = g.d_qT; a.qP = g.d_qP; a.qWork = g.d_qWork; a.qBlocks = g.qBlocks;
    return a;
}

// the grains' temperature distributions for this iteration's emission, when
// quantum heating is on for it
int launch_qheat(moc_ctx* ctx) {
    const MocDustArgs a = dust_args(ctx);
    if (!a.qActive) return MOC_OK;
    return moc_launch_qheat(ctx, a);
}

// the arguments of the dust opacity: with lgEquivalentTau in the first
// iteration, no grain counts (zero sublimation temperatures)
MocDustArgs dust_opacity_args(moc_ctx* ctx) {
    MocDustArgs a = dust_args(ctx);
    if (ctx->dust.equivTau && ctx->dust.firstIter && ctx->dust.d_tsub0) a.tsub = ctx->dust.d_tsub0;
    return a;
}

int launch_dust_opacity(moc_ctx* ctx) {
    const int rc = moc_launch_dust_opacity(ctx, dust_opacity_args(ctx));
    ctx->grid.f32_dirty = true;
    return rc;
}

int moc_opacity(moc_ctx* ctx, int ig) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->atomic.set || !g.d_ionDen) {
        ctx->err = "moc_opacity: needs an attached grid and moc_set_atomic";
        return MOC_ERR_STATE;
    }
    return run_opacity(ctx, false, nullptr);
}

MocOpacityArgs opacity_args(moc_ctx* ctx, bool skipUnchanged) {
    MocGrid& g = ctx->grid;
    MocAtomic& at = ctx->atomic;
    MocOpacityArgs a;
    a.C1 = g.nCells + 1; a.nbins = ctx->nbins; a.nSpecies = at.nSpecies;
    a.nElemUsed = at.nElemUsed; a.nstages = at.nstages; a.lgCompton = ctx->lgCompton;
    a.specElem = at.d_specElem; a.specIon = at.d_specIon; a.elemZ = at.d_elemZ;
    a.abund = at.d_abund; a.sigma = at.d_sigma; a.brems = at.d_brems; a.nuArray = ctx->d_nu;
    a.log10nu = at.d_log10nu; a.KNsigmaT = ctx->d_KNsigmaT; a.ionDen = g.d_ionDen;
    a.Ne = g.d_Ne; a.Te = g.d_Te; a.Hden = g.d_Hden;
    // with grains the gas kappa keeps a plane of its own and the total is
    // formed by the dust block, so the dust of a cell whose gas part is kept
    // is not added twice
    const bool dust = ctx->dust.set && ctx->dust.lgGas;
    a.opacity = dust ? g.d_gasOpac : g.d_opacity;
    // the mixed form's opacity in FP32 (gas runs): the plane the transport reads
    a.opacity32 = ctx->precision == MOC_PREC_MIXED && !ctx->dust.set ? g.f_opacity : nullptr;
    a.TeOld = skipUnchanged ? g.d_TeOld : nullptr;
    a.NeOld = skipUnchanged ? g.d_NeOld : nullptr;
    return a;
}

int run_opacity(moc_ctx* ctx, bool skipUnchanged, float* ms) {
    MocGrid& g = ctx->grid;
    const bool dust = ctx->dust.set && ctx->dust.lgGas;
    const MocOpacityArgs a = opacity_args(ctx, skipUnchanged);
    int rc;
    // the FP64 kernel keeps the rows of unchanged cells: bring them up to date
    if (!a.opacity32 && (rc = sync_f64_master(ctx, MOC_F_OPACITY))) return rc;
    MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    if ((rc = moc_launch_opacity(ctx, a))) return rc;
    if (a.opacity32) g.op32_master = true;
    if (dust && (rc = launch_dust_opacity(ctx))) return rc;
    MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    MOC_TRY(hipEventSynchronize(ctx->ev1), "opacity");
    if (ms) (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    g.f32_dirty = true;
    return MOC_OK;
}

int moc_upload(moc_ctx* ctx, int field, int ig, const void* host, int64_t count, int dtype,
               int layout) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_upload: no grid attached"; return MOC_ERR_STATE; }
    if (int rcs = sync_f64_master(ctx, field)) return rcs;
    double* d = field_ptr(ctx, field);
    if (!d || !host) { ctx->err = "moc_upload: unknown field"; return MOC_ERR_ARG; }
    const int64_t C1 = g.nCells + 1, B = ctx->nbins;
    const bool cb = is_cellbin(field);
    const int64_t need = field_count(ctx, field);
    if (count != need) { ctx->err = "moc_upload: element count does not match the field"; return MOC_ERR_ARG; }
    if (dtype != MOC_DT_F32 && dtype != MOC_DT_F64) { ctx->err = "moc_upload: dtype"; return MOC_ERR_ARG; }
    if (cb && layout != MOC_LAYOUT_FORTRAN && layout != MOC_LAYOUT_CELLROW) { ctx->err = "moc_upload: layout"; return MOC_ERR_ARG; }
    std::vector<double> buf((size_t)need);
    for (int64_t i = 0; i < need; ++i)
        buf[i] = dtype == MOC_DT_F64 ? ((const double*)host)[i] : (double)((const float*)host)[i];
    if (cb && layout != device_layout(field)) {  // transpose (cell,bin) <-> (bin,cell)
        std::vector<double> t((size_t)need);
        if (layout == MOC_LAYOUT_FORTRAN)   // host cell-fastest -> device bin-fastest
            for (int64_t b = 0; b < B; ++b)
                for (int64_t c = 0; c < C1; ++c) t[c * B + b] = buf[b * C1 + c];
        else
            for (int64_t c = 0; c < C1; ++c)
                for (int64_t b = 0; b < B; ++b) t[b * C1 + c] = buf[c * B + b];
        buf.swap(t);
    }
    int rc = h2d(ctx, d, buf.data(), (size_t)need * 8, "upload");
    if (rc) return rc;
    g.f32_dirty = true;
    return MOC_OK;
}

int moc_download(moc_ctx* ctx, int field, int ig, void* host, int64_t count, int dtype,
                 int layout) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_download: no grid attached"; return MOC_ERR_STATE; }
    if (field == MOC_F_CONVERGED || field == MOC_F_BLACK) {
        const int* d = field == MOC_F_CONVERGED ? g.d_conv : g.d_black;
        const int64_t C1 = g.nCells + 1;
        if (!d) { ctx->err = "moc_download: no balance run yet"; return MOC_ERR_STATE; }
        if (count != C1 || !host) { ctx->err = "moc_download: element count does not match the field"; return MOC_ERR_ARG; }
        std::vector<int> buf((size_t)C1);
        MOC_TRY(hipMemcpyAsync(buf.data(), d, (size_t)C1 * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
        MOC_TRY(hipStreamSynchronize(ctx->stream), "download");
        for (int64_t i = 0; i < C1; ++i) {
            if (dtype == MOC_DT_I32) ((int*)host)[i] = buf[i];
            else if (dtype == MOC_DT_F64) ((double*)host)[i] = buf[i];
            else ((float*)host)[i] = (float)buf[i];
        }
        return MOC_OK;
    }
    if (int rcs = sync_f64_master(ctx, field)) return rcs;
    double* d = field_ptr(ctx, field);
    if (!d || !host) { ctx->err = "moc_download: unknown field"; return MOC_ERR_ARG; }
    const int64_t C1 = g.nCells + 1, B = ctx->nbins;
    const bool cb = is_cellbin(field);
    const int64_t need = field_count(ctx, field);
    if (count != need) { ctx->err = "moc_download: element count does not match the field"; return MOC_ERR_ARG; }
    if (dtype != MOC_DT_F32 && dtype != MOC_DT_F64) { ctx->err = "moc_download: dtype"; return MOC_ERR_ARG; }
    std::vector<double> buf((size_t)need);
    MOC_TRY(hipMemcpyAsync(buf.data(), d, (size_t)need * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "download");
    if (cb && layout != device_layout(field)) {
        std::vector<double> t((size_t)need);
        if (layout == MOC_LAYOUT_FORTRAN)   // device bin-fastest -> host cell-fastest
            for (int64_t c = 0; c < C1; ++c)
                for (int64_t b = 0; b < B; ++b) t[b * C1 + c] = buf[c * B + b];
        else
            for (int64_t b = 0; b < B; ++b)
                for (int64_t c = 0; c < C1; ++c) t[c * B + b] = buf[b * C1 + c];
        buf.swap(t);
    }
    for (int64_t i = 0; i < need; ++i) {
        if (dtype == MOC_DT_F64) ((double*)host)[i] = buf[i];
        else ((float*)host)[i] = (float)buf[i];
    }
    return MOC_OK;
}

int moc_device_ptr(moc_ctx* ctx, int field, int ig, void** dptr, int64_t* count, int* dtype) {
    if (!ctx || !dptr) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_device_ptr: no grid attached"; return MOC_ERR_STATE; }
    if (int rcs = sync_f64_master(ctx, field)) return rcs;
    double* d = field_ptr(ctx, field);
    if (!d) { ctx->err = "moc_device_ptr: unknown field"; return MOC_ERR_ARG; }
    *dptr = d;
    if (count) *count = field_count(ctx, field);
    if (dtype) *dtype = MOC_DT_F64;
    g.f32_dirty = true;  // the caller may write through the pointer
    return MOC_OK;
}

int moc_reset_tallies(moc_ctx* ctx, int ig) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_reset_tallies: no grid attached"; return MOC_ERR_STATE; }
    const size_t n = ((size_t)g.nCells + 1) * ctx->nbins;
    MOC_TRY(hipMemsetAsync(g.d_J, 0, n * 8, ctx->stream), "reset J");
    MOC_TRY(hipMemsetAsync(g.d_escaped, 0, n * 8, ctx->stream), "reset escaped");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "reset");
    return MOC_OK;
}

int moc_transport(moc_ctx* ctx, int iStar, int64_t n, uint64_t seed, uint32_t ctr_base,
                  moc_transport_stats* stats) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached) { ctx->err = "moc_transport: no grid attached"; return MOC_ERR_STATE; }
    if (iStar < 1 || iStar > ctx->nStars || n < 0) { ctx->err = "moc_transport: iStar/n"; return MOC_ERR_ARG; }
    return run_transport(ctx, iStar, n, seed, ctr_base, stats);
}

}  // extern "C"

namespace {

// One transport launch: star iStar's n packets, or (iStar = 0, a batch of
// grids) every star's nPhotons packets, each in its own grid.
// transport arguments of source s (0-based) with n packets; the tuning of the
// context; the batch views when the context holds a batch of grids
MocTransportArgs<double> transport_args_f64(moc_ctx* ctx, int s, int64_t n, uint64_t seed,
                                            uint32_t ctr_base, double deltaE) {
    MocGrid& g = ctx->grid;
    MocTransportArgs<double> a;
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nCells = g.nCells; a.sym = g.sym; a.plane = g.plane;
    a.ax = g.d_ax; a.ay = g.d_ay; a.az = g.d_az; a.vx = g.d_vx; a.vy = g.d_vy; a.vz = g.d_vz;
    a.active = g.d_active; a.dense = g.dense ? 1 : 0; a.R_out = g.R_out;
    a.nbins = ctx->nbins; a.nAngles = ctx->nAngles; a.lgCompton = ctx->lgCompton;
    a.ionEdge1 = ctx->ionEdge1; a.nuArray = ctx->d_nu; a.KNsigmaT = ctx->d_KNsigmaT;
    a.KNcdf = ctx->d_KNcdf; a.Pcomp = ctx->d_Pcomp;
    a.sourceCDF = ctx->d_sourceCDF + (size_t)s * ctx->nbins;
    a.opacity = g.d_opacity; a.Ne = g.d_Ne; a.recPDF = g.d_recPDF; a.totalLines = g.d_totalLines;
    a.scaOpac = g.d_scaOpac; a.lgDust = ctx->dust.set ? 1 : 0;
    a.dustOnly = ctx->dust.set && !ctx->dust.lgGas ? 1 : 0;
    a.J = g.d_J; a.escaped = g.d_escaped; a.Qphot = ctx->d_Qphot; a.counters = ctx->d_counters;
    for (int k = 0; k < 3; ++k) { a.star[k] = ctx->starPos[3 * s + k]; a.starIdx[k] = ctx->starIdx[3 * s + k] - 1; }
    a.deltaE = deltaE; a.n = n; a.seed = seed; a.ctr_base = ctr_base; a.ctrDev = nullptr;
    a.emitBatch = ctx->tune.emitBatch; a.segsPerTrip = ctx->tune.segsPerTrip;
    a.nGrids = g.nGrids > 1 ? g.nGrids : 0; a.grids = g.d_view64; a.gridFirst = g.d_gridFirst;
    a.nested = g.nested; a.nestStart = g.nestStart;
    a.mir[0] = g.ax.empty() ? 0.0 : g.ax[0]; a.mir[1] = g.ay.empty() ? 0.0 : g.ay[0];
    a.mir[2] = g.az.empty() ? 0.0 : g.az[0];
    return a;
}

// The hot block of the mixed transport's J tallies (single grid): a box of
// h x h x h cells around star 1 -- the cells every stellar packet crosses,
// whose slot the kernel computes from the cell indices -- times as many bins
// as the LDS of a workgroup of moc_transport_hot_block() lanes then holds (at
// most kHotBins); the bins are chosen on the device after each launch
// (k_hot_binsum, k_hot_select).  The partial buffer has a row per workgroup of
// the launch at the current wave count.  Allocates, so it is called before
// any graph capture.
constexpr int kHotBins = 256;
// workgroups (and work rows of nTbins^2 doubles) of the quantum-heating solver
constexpr int kQHeatBlocks = 512;
int ensure_hot(moc_ctx* ctx) {
    MocGrid& g = ctx->grid;
    const int B = ctx->nbins;
    if (ctx->precision != MOC_PREC_MIXED || !ctx->tune.hotJ || g.nGrids > 1 || g.plane || B > 8192 ||
        g.nCells < 1)
        return MOC_OK;
    const int blk = moc_transport_hot_block();
    const int blocksNeed = (int)((64LL * ctx->tune.transportWaves + blk - 1) / blk);
    if (g.hotTried && g.nHotC == 0) return MOC_OK;  // no hot block fits this grid
    int rc;
    if (!g.hotTried) {
        g.hotTried = true;
        int dev = 0, maxLds = 0;
        MOC_TRY(hipGetDevice(&dev), "device");
        MOC_TRY(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "LDS size");
        const size_t cap = std::min<size_t>((size_t)maxLds, 160 * 1024);
        const size_t geo = moc_transport_hot_lds(g.nx, g.ny, g.nz, 0) + 1024;
        if (cap <= geo) return MOC_OK;
        const int slots = (int)((cap - geo) / sizeof(float));
        // the largest cube that leaves at least 128 bins (or all of them)
        const int nbMin = std::min(B, 128);
        int h = 1;
        while ((h + 1) * (h + 1) * (h + 1) * nbMin <= slots) ++h;
        const int n[3] = {g.nx, g.ny, g.nz};
        int lo[3], hn[3];
        for (int k = 0; k < 3; ++k) {
            const int si = ctx->starIdx[k] - 1;  // 0-based index of the star's cell
            hn[k] = std::min(h, n[k]);
            lo[k] = g.sym ? 0 : std::max(0, std::min(si - hn[k] / 2, n[k] - hn[k]));
        }
        const int nc = hn[0] * hn[1] * hn[2];
        const int nb = std::min({B, kHotBins, slots / nc});
        if (nc < 8 || nb < 1) return MOC_OK;
        std::vector<int> cellList((size_t)nc, 0);
        for (int x = 0; x < hn[0]; ++x)
            for (int y = 0; y < hn[1]; ++y)
                for (int z = 0; z < hn[2]; ++z) {
                    const int ix = lo[0] + x, iy = lo[1] + y, iz = lo[2] + z;
                    cellList[(size_t)(x * hn[1] + y) * hn[2] + z] =
                        g.dense ? 1 + iz + g.nz * (iy + g.ny * ix)
                                : g.active[(size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz)];
                }
        std::vector<int> binMap((size_t)B, -1), binList((size_t)nb, -1);
        if ((rc = dalloc_n(ctx, &g.d_hotCellList, (int64_t)nc, "hot cells")) ||
            (rc = dalloc_n(ctx, &g.d_hotBin, (int64_t)B, "hot bins")) ||
            (rc = dalloc_n(ctx, &g.d_hotBinList, (int64_t)nb, "hot bins")) ||
            (rc = dalloc_n(ctx, &g.d_hotS, (int64_t)B, "hot bin sums")))
            return rc;
        if ((rc = h2d(ctx, g.d_hotCellList, cellList.data(), cellList.size() * sizeof(int), "hot cells")) ||
            (rc = h2d(ctx, g.d_hotBin, binMap.data(), binMap.size() * sizeof(int), "hot bins")) ||
            (rc = h2d(ctx, g.d_hotBinList, binList.data(), binList.size() * sizeof(int), "hot bins")))
            return rc;
        for (int k = 0; k < 3; ++k) { g.hotLo[k] = lo[k]; g.hotN[k] = hn[k]; }
        g.nHotC = nc;
        g.nHotB = nb;
    }
    if (g.hotBlocks < blocksNeed) {
        dfree(g.d_hotPart);
        g.hotBlocks = 0;
        if ((rc = dalloc_n(ctx, &g.d_hotPart, (int64_t)blocksNeed * g.nHotC * g.nHotB, "hot partials")))
            return rc;
        g.hotBlocks = blocksNeed;
    }
    return MOC_OK;
}

MocTransportArgs<float> transport_args_f32(moc_ctx* ctx, int s, int64_t n, uint64_t seed,
                                           uint32_t ctr_base, double deltaE) {
    MocGrid& g = ctx->grid;
    MocTransportArgs<float> a;
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nCells = g.nCells; a.sym = g.sym; a.plane = g.plane;
    a.ax = g.f_ax; a.ay = g.f_ay; a.az = g.f_az; a.vx = g.f_vx; a.vy = g.f_vy; a.vz = g.f_vz;
    a.active = g.d_active; a.dense = g.dense ? 1 : 0; a.R_out = (float)g.R_out;
    a.nbins = ctx->nbins; a.nAngles = ctx->nAngles; a.lgCompton = ctx->lgCompton;
    a.ionEdge1 = (float)ctx->ionEdge1; a.nuArray = ctx->f_nu; a.KNsigmaT = ctx->f_KNsigmaT;
    a.KNcdf = ctx->f_KNcdf; a.Pcomp = ctx->f_Pcomp;
    a.sourceCDF = ctx->f_sourceCDF + (size_t)s * ctx->nbins;
    a.opacity = g.f_opacity; a.Ne = g.f_Ne; a.recPDF = g.f_recPDF; a.totalLines = g.f_totalLines;
    a.scaOpac = g.f_scaOpac; a.lgDust = ctx->dust.set ? 1 : 0;
    a.dustOnly = ctx->dust.set && !ctx->dust.lgGas ? 1 : 0;
    a.J = g.d_J; a.escaped = g.d_escaped; a.Qphot = ctx->d_Qphot; a.counters = ctx->d_counters;
    for (int k = 0; k < 3; ++k) { a.star[k] = (float)ctx->starPos[3 * s + k]; a.starIdx[k] = ctx->starIdx[3 * s + k] - 1; }
    a.deltaE = (float)deltaE; a.n = n; a.seed = seed; a.ctr_base = ctr_base; a.ctrDev = nullptr;
    a.emitBatch = ctx->tune.emitBatch; a.segsPerTrip = ctx->tune.segsPerTrip;
    a.nGrids = g.nGrids > 1 ? g.nGrids : 0; a.grids = g.d_view32; a.gridFirst = g.d_gridFirst;
    a.nested = g.nested; a.nestStart = g.nestStart;
    a.mir[0] = g.ax.empty() ? 0.f : (float)g.ax[0]; a.mir[1] = g.ay.empty() ? 0.f : (float)g.ay[0];
    a.mir[2] = g.az.empty() ? 0.f : (float)g.az[0];
    const bool hot = ctx->tune.hotJ && a.nGrids == 0 && g.nHotC > 0 && g.hotBlocks > 0;
    a.nHotC = hot ? g.nHotC : 0; a.nHotB = hot ? g.nHotB : 0;
    for (int k = 0; k < 3; ++k) { a.hotLo[k] = g.hotLo[k]; a.hotN[k] = g.hotN[k]; }
    a.hotBin = g.d_hotBin; a.hotCellList = g.d_hotCellList; a.hotS = g.d_hotS;
    a.hotBinList = g.d_hotBinList; a.hotPart = g.d_hotPart; a.hotBlocks = g.hotBlocks;
    return a;
}

// a batch of grids: per-grid packet ranges and views (star k lights grid k),
// uploaded to the device; *n becomes the packet count of the whole batch
int upload_batch_views(moc_ctx* ctx, int iStar, int64_t* n) {
    MocGrid& g = ctx->grid;
    const int s = iStar > 0 ? iStar - 1 : 0;
    std::vector<int64_t> first;
    std::vector<MocGridView<double>> v64;
    std::vector<MocGridView<float>> v32;
    {
        const int nG = g.nGrids;
        first.assign((size_t)nG + 1, 0);
        for (int k = 0; k < nG; ++k) {
            const int64_t cnt = iStar == 0 ? (int64_t)ctx->nPhotons[k] : (k == s ? *n : 0);
            if (cnt < 0) { ctx->err = "moc_transport: negative packet count"; return MOC_ERR_ARG; }
            first[k + 1] = first[k] + cnt;
            if (cnt > 0xffffffffLL) { ctx->err = "moc_transport: more than 2^32 packets in one grid"; return MOC_ERR_ARG; }
        }
        *n = first[nG];
        v64.resize(nG);
        v32.resize(nG);
        for (int k = 0; k < nG; ++k) {
            const double dE = ctx->nPhotons[k] > 0 ? ctx->Lstar[k] / ctx->nPhotons[k] : 0.0;
            if (!(dE >= 0.0)) { ctx->err = "moc_transport: deltaEUsed < 0"; return MOC_ERR_ARG; }
            MocGridView<double>& a = v64[k];
            MocGridView<float>& b = v32[k];
            a.ax = g.d_ax + g.goffx[k]; a.ay = g.d_ay + g.goffy[k]; a.az = g.d_az + g.goffz[k];
            a.vx = g.d_vx + g.goffx[k]; a.vy = g.d_vy + g.goffy[k]; a.vz = g.d_vz + g.goffz[k];
            b.ax = g.f_ax + g.goffx[k]; b.ay = g.f_ay + g.goffy[k]; b.az = g.f_az + g.goffz[k];
            b.vx = g.f_vx + g.goffx[k]; b.vy = g.f_vy + g.goffy[k]; b.vz = g.f_vz + g.goffz[k];
            a.active = b.active = g.gdense[k] ? nullptr : g.d_active + g.goffa[k];
            a.sourceCDF = ctx->d_sourceCDF + (size_t)k * ctx->nbins;
            b.sourceCDF = ctx->f_sourceCDF + (size_t)k * ctx->nbins;
            a.nx = b.nx = g.gnx[k]; a.ny = b.ny = g.gny[k]; a.nz = b.nz = g.gnz[k];
            a.base = b.base = g.gbase[k];
            a.sym = b.sym = g.sym;
            a.plane = b.plane = g.plane;
            for (int d = 0; d < 3; ++d) {
                a.starIdx[d] = b.starIdx[d] = ctx->starIdx[3 * k + d] - 1;
                a.star[d] = ctx->starPos[3 * k + d];
                b.star[d] = (float)ctx->starPos[3 * k + d];
            }
            a.R_out = g.gR_out[k]; b.R_out = (float)g.gR_out[k];
            a.deltaE = dE; b.deltaE = (float)dE;
        }
        int rc;
        if ((rc = h2d(ctx, g.d_gridFirst, first.data(), first.size() * 8, "grid packets")) ||
            (rc = h2d(ctx, g.d_view64, v64.data(), v64.size() * sizeof(MocGridView<double>), "grid views")) ||
            (rc = h2d(ctx, g.d_view32, v32.data(), v32.size() * sizeof(MocGridView<float>), "grid views")))
            return rc;
    }
    return MOC_OK;
}

int run_transport(moc_ctx* ctx, int iStar, int64_t n, uint64_t seed, uint32_t ctr_base,
                  moc_transport_stats* stats) {
    MocGrid& g = ctx->grid;
    const int s = iStar > 0 ? iStar - 1 : 0;
    const double deltaE = ctx->Lstar[s] / ctx->nPhotons[s];  // photon_mod.f90:88
    if (!(deltaE >= 0.0)) { ctx->err = "moc_transport: deltaEUsed < 0"; return MOC_ERR_ARG; }
    int rc;
    if (g.nested && (rc = upload_nested_views(ctx, s))) return rc;
    if (g.batch() && (rc = upload_batch_views(ctx, iStar, &n))) return rc;
    MOC_TRY(hipMemsetAsync(ctx->d_counters, 0, MOC_CNT_N * sizeof(unsigned long long), ctx->stream), "counters");
    MOC_TRY(hipMemsetAsync(ctx->d_Qphot, 0, sizeof(double), ctx->stream), "Qphot");
    float ms = 0.f;
    if (ctx->precision == MOC_PREC_FP64) {
        MocTransportArgs<double> a = transport_args_f64(ctx, s, n, seed, ctr_base, deltaE);
        rc = moc_launch_transport_f64(ctx, a, &ms);
    } else {
        if ((rc = refresh_f32(ctx)) || (rc = ensure_hot(ctx))) return rc;
        MocTransportArgs<float> a = transport_args_f32(ctx, s, n, seed, ctr_base, deltaE);
        rc = moc_launch_transport_f32(ctx, a, &ms);
    }
    if (rc) return rc;
    unsigned long long cnt[MOC_CNT_N];
    double q = 0.0;
    MOC_TRY(hipMemcpyAsync(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream), "counters");
    MOC_TRY(hipMemcpyAsync(&q, ctx->d_Qphot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "Qphot");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "counters");
    if (stats) {
        stats->Qphot = q;
        stats->segments = (double)cnt[MOC_CNT_SEGMENTS];
        stats->emissions = (double)cnt[MOC_CNT_EMISSIONS];
        stats->absorbed = (double)cnt[MOC_CNT_ABSORBED];
        stats->scattered = (double)cnt[MOC_CNT_SCATTERED];
        stats->escaped = (double)cnt[MOC_CNT_ESCAPED];
        stats->lines = (double)cnt[MOC_CNT_LINES];
        stats->lost = (double)cnt[MOC_CNT_LOST];
        stats->trapped = (double)cnt[MOC_CNT_TRAPPED];
        stats->kernel_ms = ms;
    }
    if (cnt[MOC_CNT_ERROR]) {
        ctx->err = "moc_transport: " + std::to_string(cnt[MOC_CNT_ERROR]) +
                   " packets reached an insane state (index outside the grid or emission "
                   "from an inactive cell)";
        return MOC_ERR_KERNEL;
    }
    return MOC_OK;
}

}  // namespace

extern "C" {

int moc_set_balance(moc_ctx* ctx, const moc_balance_tables* t) {
    if (!ctx) return MOC_ERR_ARG;
    drop_graph(ctx);
    if (!ctx->atomic.set) { ctx->err = "moc_set_balance: call moc_set_atomic first"; return MOC_ERR_STATE; }
    MocAtomic& at = ctx->atomic;
    if (!t || t->maxLev < 1 || t->maxLev > MOC_MAXLEV || t->nShells < 0 || !t->speciesShell ||
        at.nElemUsed > MOC_MAXE || at.nstages > MOC_MAXS || t->nAtoms < 0) {
        ctx->err = "moc_set_balance: invalid tables (maxLev <= 27, <= 30 elements, <= 31 stages)";
        return MOC_ERR_ARG;
    }
    bool hasH = false;
    for (int e = 0; e < at.nElemUsed; ++e) hasH |= at.elemZ[e] == 1;
    if (!hasH) { ctx->err = "moc_set_balance: hydrogen must be switched on"; return MOC_ERR_ARG; }
    for (int sp = 0; sp < at.nSpecies; ++sp)
        if (t->speciesShell[sp] > t->speciesShell[sp + 1] || t->speciesShell[sp + 1] > t->nShells) {
            ctx->err = "moc_set_balance: speciesShell ranges";
            return MOC_ERR_ARG;
        }
    for (int k = 0; k < t->nAtoms; ++k)
        if (t->atomMeta[8 * k] < 1 || t->atomMeta[8 * k] > 64 || t->atomMeta[8 * k + 1] < 1 ||
            t->atomMeta[8 * k + 1] > 64) {
            ctx->err = "moc_set_balance: n-level ion sizes";
            return MOC_ERR_ARG;
        }
    free_balance(ctx->balance);
    free_emission(ctx->emission);  // its n-level ion list refers to these atoms
    MocBalance& b = ctx->balance;
    MocBalanceArgs& a = b.a;
    const int B = ctx->nbins, S = at.nSpecies;
    b.ncol = 3 * t->nShells + 5;
    // host-side index tables
    std::vector<int> slot(31, -1), specOf((size_t)at.nElemUsed * 32, -1), chexIdx(31 * 32, -1);
    for (int e = 0; e < at.nElemUsed; ++e) slot[at.elemZ[e]] = e;
    for (int sp = 0; sp < S; ++sp) specOf[(size_t)at.specElem[sp] * 32 + at.specIon[sp]] = sp;
    for (int k = 0; k < t->nChex; ++k) chexIdx[t->chexZ[k] * 32 + t->chexIon[k]] = k;  // last row wins
    int rc;
    const int* p_i;
    const double* p_d;
    const int64_t* p_l;
#define PUT(dst, src, n, what) do { if ((rc = bal_put(ctx, src, (size_t)(n), &dst, what))) return rc; } while (0)
    PUT(a.elemSlot, slot.data(), 31, "elemSlot");
    PUT(a.specOf, specOf.data(), specOf.size(), "specOf");
    PUT(a.specShell, t->speciesShell, S + 1, "specShell");
    PUT(a.shellNum, t->shellNum, t->nShells, "shellNum");
    PUT(a.outShell, t->outShell, 31 * 32, "outShell");
    PUT(a.statW, t->statW0, 31 * 32, "statW0");
    PUT(a.statW1, t->statW1, 31 * 32, "statW1");
    PUT(a.rrb, t->rrb, 6 * 31 * 31, "rrb");
    PUT(a.drb_c, t->drb_c, 9 * 31 * 31, "drb_c");
    PUT(a.drb_e, t->drb_e, 9 * 31 * 31, "drb_e");
    PUT(a.rrec, t->rrec, 2 * 31 * 31, "rrec");
    PUT(a.rnew, t->rnew, 4 * 31 * 31, "rnew");
    PUT(a.diel, t->diel, 8 * (size_t)t->nDiel, "diel");
    PUT(a.aldro, t->aldro, 6 * (size_t)t->nAldro, "aldro");
    PUT(a.cion_cf, t->cionCF, 5 * 31 * 31, "cionCF");
    PUT(a.cion_th, t->cionTherm, 31 * 32, "cionTherm");
    PUT(a.nauger, t->nAuger, 31 * 32 * 11, "nAuger");
    PUT(a.auger, t->auger, 31 * 32 * 11 * 11, "auger");
    {
        // per-shell view of the Auger and reverse charge-exchange tables
        // (ionBalance :1826-1929): the chains then index by shell alone
        std::vector<int> info((size_t)std::max(t->nShells, 1), 0), aug((size_t)std::max(t->nShells, 1), 0);
        for (int sp = 0; sp < S; ++sp) {
            const int z = at.elemZ[at.specElem[sp]], ion = at.specIon[sp];
            for (int sh = t->speciesShell[sp]; sh < t->speciesShell[sp + 1]; ++sh) {
                const int shn = t->shellNum[sh];
                if (shn < 0 || shn > 10 || ion < 1 || ion > 31 || z - ion + 1 < 0) {
                    ctx->err = "moc_set_balance: shell number or stage out of range";
                    return MOC_ERR_ARG;
                }
                const int na = t->nAuger[(z * 32 + ion) * 11 + shn];
                if (na < 0 || na > 10) {
                    ctx->err = "moc_set_balance: Auger electron count above 10";
                    return MOC_ERR_ARG;
                }
                const bool cx = shn == t->outShell[z * 32 + (z - ion + 1)] && ion == 1 && (z == 7 || z == 8);
                info[sh] = na | ((cx ? z - 6 : 0) << 8);
                aug[sh] = ((z * 32 + ion) * 11 + shn) * 11;
            }
        }
        PUT(a.shInfo, info.data(), info.size(), "shInfo");
        PUT(a.shAug, aug.data(), aug.size(), "shAug");
        // The rates out of each stage of the heavy-element chains (ionBalance
        // :1826-1929) as gather lists: the sequential walk over (stage, shell,
        // Auger electrons) adds its terms to out(target stage); here each
        // target stage (a task) keeps its own terms in the walk's order, so
        // that a lane per task forms the same sums, addition for addition.
        // Term code: kind (0 collisional c1 Ne, 1 the shell's rate, 2 an
        // Auger term) | reverse charge-exchange slot << 2 | source stage << 4
        // | shell << 9; the Auger yield of kind-2 terms in htAug.  Tasks are
        // sorted by length (longest first) so that the lanes share the work.
        const int NS = at.nstages;
        std::vector<std::vector<int>> code;
        std::vector<std::vector<double>> yld;
        std::vector<int> tSlot, tMx;
        for (int e = 0; e < at.nElemUsed; ++e) {
            const int z = at.elemZ[e];
            if (z < 3) continue;
            const int nIon = std::min(z, NS - 1), base = (int)code.size();
            for (int mx = 1; mx <= nIon; ++mx) {
                tSlot.push_back(e);
                tMx.push_back(mx);
                code.emplace_back();
                yld.emplace_back();
            }
            for (int ion = 1; ion <= nIon; ++ion) {
                code[base + ion - 1].push_back(0 | (ion << 4));
                yld[base + ion - 1].push_back(0.);
                const int spec = specOf[(size_t)e * 32 + ion];
                if (spec < 0) continue;
                for (int sh = t->speciesShell[spec]; sh < t->speciesShell[spec + 1]; ++sh) {
                    const int na = info[sh] & 0xff, cx = info[sh] >> 8;
                    code[base + ion - 1].push_back(1 | (cx << 2) | (ion << 4) | (sh << 9));
                    yld[base + ion - 1].push_back(0.);
                    for (int ne = 2; ne <= na; ++ne) {
                        const int mx = std::min(ion + ne - 1, nIon);
                        code[base + mx - 1].push_back(2 | (cx << 2) | (ion << 4) | (sh << 9));
                        yld[base + mx - 1].push_back(t->auger[aug[sh] + ne]);
                    }
                }
            }
        }
        std::vector<int> order(code.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return code[x].size() > code[y].size(); });
        std::vector<int> hs, hm, hb{0}, hc;
        std::vector<double> hy;
        for (int k : order) {
            hs.push_back(tSlot[k]);
            hm.push_back(tMx[k]);
            hc.insert(hc.end(), code[k].begin(), code[k].end());
            hy.insert(hy.end(), yld[k].begin(), yld[k].end());
            hb.push_back((int)hc.size());
        }
        a.nHTask = (int)hs.size();
        PUT(a.htSlot, hs.data(), hs.size(), "chain tasks");
        PUT(a.htMx, hm.data(), hm.size(), "chain tasks");
        PUT(a.htBeg, hb.data(), hb.size(), "chain tasks");
        PUT(a.htCode, hc.data(), hc.size(), "chain terms");
        PUT(a.htAug, hy.data(), hy.size(), "chain terms");
        // the charge-exchange coefficient of the heavy chains, unscaled (ionBalance
        // leaves it at its 1e4 K value): chexFit(0) 1e-9 of the (Z, stage) row
        std::vector<double> cu(31 * 32, 0.);
        for (int k = 0; k < 31 * 32; ++k)
            if (chexIdx[k] >= 0) cu[k] = t->chexFit[4 * (size_t)chexIdx[k]] * 1.e-9;
        PUT(a.chexU, cu.data(), cu.size(), "chexU");
        // dielectronic rows (dielectronic :3204-3305, the last matching row
        // wins): per (Z, stage) the Nussbaumer & Storey row in force below and
        // at or above 2e4 K, and the Aldrovandi & Pequignot row
        std::vector<int> di(31 * 32 * 2, -1), ai(31 * 32, -1);
        for (int i = 0; i < t->nDiel; ++i) {
            const double* r = t->diel + 8 * (size_t)i;
            const int z = (int)r[0], ion = (int)r[0] + 1 - (int)r[1], g = (int)r[7];
            if (z < 0 || z > 30 || ion < 0 || ion > 31) continue;
            if (g == 0 || g == 1) di[(z * 32 + ion) * 2 + 0] = i;
            if (g == 0 || g == 2) di[(z * 32 + ion) * 2 + 1] = i;
        }
        for (int i = 0; i < t->nAldro; ++i) {
            const double* r = t->aldro + 6 * (size_t)i;
            const int z = (int)r[0], ion = (int)r[0] - (int)r[1] + 1;
            if (z < 0 || z > 30 || ion < 0 || ion > 31) continue;
            ai[z * 32 + ion] = i;
        }
        PUT(a.dielIdx, di.data(), di.size(), "dielIdx");
        PUT(a.aldroIdx, ai.data(), ai.size(), "aldroIdx");
    }
    PUT(a.chexIdx, chexIdx.data(), 31 * 32, "chexIdx");
    PUT(a.chexTab, t->chexFit, 4 * (size_t)t->nChex, "chexFit");
    PUT(a.atomIdx, t->atomIdx, 31 * 32, "atomIdx");
    {
        // the ions whose n-level cooling thermBalance sums, in its order (:1065-1093)
        std::vector<int> ns, ni, na;
        for (int e = 0; e < at.nElemUsed; ++e) {
            const int z = at.elemZ[e];
            if (z < 3) continue;
            for (int ion = 1; ion <= std::min(z + 1, at.nstages); ++ion) {
                const int k = t->atomIdx[z * 32 + ion];
                if (k < 0) continue;
                ns.push_back(e); ni.push_back(ion); na.push_back(k);
            }
        }
        a.nNlv = (int)na.size();
        PUT(a.nlvSlot, ns.data(), ns.size(), "nlvSlot");
        PUT(a.nlvIon, ni.data(), ni.size(), "nlvIon");
        PUT(a.nlvAtom, na.data(), na.size(), "nlvAtom");
        // ions of more than 5 levels (as cut at maxLev) are solved by teams of
        // 8, 16 or 32 lanes, in the level classes of nlevel_team_class
        std::vector<int> tl[kNlevTeams];
        for (int k = 0; k < (int)na.size(); ++k) {
            const int c = nlevel_team_class((int)std::min<int64_t>(t->atomMeta[8 * na[k]], t->maxLev));
            if (c >= 0) tl[c].push_back(k);
        }
        for (int c = 0; c < kNlevTeams; ++c) {
            a.nNlvT[c] = (int)tl[c].size();
            PUT(a.nlvT[c], tl[c].data(), tl[c].size(), "nlvT");
        }
        // level counts as solved, for the emission's line lists
        b.atomLevels.resize((size_t)t->nAtoms);
        for (int k = 0; k < t->nAtoms; ++k)
            b.atomLevels[k] = (int)std::min<int64_t>(t->atomMeta[8 * k], t->maxLev);
    }
    b.atomIdx.assign(t->atomIdx, t->atomIdx + 31 * 32);
    PUT(a.atomMeta, t->atomMeta, 8 * (size_t)t->nAtoms, "atomMeta");
    PUT(a.atomAtp, t->atomAtp, t->nAtoms, "atomAtp");
    PUT(a.atomLogT, t->atomLogT, t->nLogT, "atomLogT");
    PUT(a.atomG, t->atomG, t->nLev, "atomG");
    PUT(a.atomE, t->atomE, t->nLev, "atomE");
    PUT(a.atomA, t->atomA, t->nA, "atomA");
    PUT(a.atomQom, t->atomQom, t->nQ, "atomQom");
    PUT(a.atomQ2, t->atomQ2, t->nQ, "atomQ2");
    if (t->atomRecKind) {
        PUT(a.atomRecKind, t->atomRecKind, t->nAtoms, "atomRecKind");
        PUT(a.atomRec, t->atomRec, 11 * (size_t)t->nLev, "atomRec");
        PUT(a.atomRecValid, t->atomRecValid, t->nLev, "atomRecValid");
    } else {
        std::vector<int> none((size_t)t->nAtoms, 0);
        PUT(a.atomRecKind, none.data(), none.size(), "atomRecKind");
        a.atomRec = nullptr;
        a.atomRecValid = nullptr;
    }
    {
        // the rate weights in band form: columns sorted by the bins where they
        // are nonzero (by first bin, last bin or both, whichever leaves the
        // least work), each column tile contracted over its own bins only
        const int ncol = b.ncol, TC = moc_rate_tile_cols();
        std::vector<int> lo((size_t)ncol, B), hi((size_t)ncol, -1);
        for (int bb = 0; bb < B; ++bb)
            for (int c = 0; c < ncol; ++c)
                if (t->W[(size_t)bb * ncol + c] != 0.0) {
                    if (lo[c] > bb) lo[c] = bb;
                    hi[c] = bb;
                }
        std::vector<int> best, bestLo, bestHi;
        long long bestWork = -1;
        for (int key = 0; key < 3; ++key) {
            std::vector<int> perm((size_t)ncol);
            for (int c = 0; c < ncol; ++c) perm[c] = c;
            std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
                if (key == 0) return lo[x] < lo[y];
                if (key == 1) return hi[x] < hi[y];
                return lo[x] != lo[y] ? lo[x] < lo[y] : hi[x] < hi[y];
            });
            const int nt = (ncol + TC - 1) / TC;
            std::vector<int> tlo((size_t)nt, B), thi((size_t)nt, -1);
            long long work = 0;
            for (int tt = 0; tt < nt; ++tt) {
                for (int n = tt * TC; n < std::min(ncol, (tt + 1) * TC); ++n) {
                    tlo[tt] = std::min(tlo[tt], lo[perm[n]]);
                    thi[tt] = std::max(thi[tt], hi[perm[n]]);
                }
                work += std::max(0, thi[tt] + 1 - tlo[tt]);
            }
            if (bestWork < 0 || work < bestWork) {
                bestWork = work;
                best.swap(perm);
                bestLo.swap(tlo);
                bestHi.swap(thi);
            }
        }
        std::vector<double> Wp((size_t)B * ncol);
        for (int bb = 0; bb < B; ++bb)
            for (int n = 0; n < ncol; ++n) Wp[(size_t)bb * ncol + n] = t->W[(size_t)bb * ncol + best[n]];
        PUT(p_d, Wp.data(), Wp.size(), "W");
        PUT(a.ratePerm, best.data(), best.size(), "rate column order");
        PUT(a.rateLo, bestLo.data(), bestLo.size(), "rate tile bins");
        PUT(a.rateHi, bestHi.data(), bestHi.size(), "rate tile bins");
    }
    b.d_W = nullptr;  // W lives in b.owned
#undef PUT
    (void)p_i; (void)p_l;
    if ((rc = dalloc_n(ctx, &b.d_counters, 4, "balance counters"))) return rc;
    a.elemZ = at.d_elemZ;
    a.abund = at.d_abund;
    a.specElem = at.d_specElem;
    a.specIon = at.d_specIon;
    a.NE = at.nElemUsed;
    a.NS = at.nstages;
    a.nX = 0;  // packed stage rows: stages 0 .. min(Z+1, NS) of each element
    for (int e = 0; e < at.nElemUsed; ++e) {
        a.xoff[e] = a.nX;
        a.nX += std::min(at.elemZ[e] + 1, at.nstages) + 1;
    }
    a.xoff[at.nElemUsed] = a.nX;
    a.nSpec = S;
    a.nShells = t->nShells;
    a.nDiel = t->nDiel;
    a.nAldro = t->nAldro;
    a.maxLev = t->maxLev;
    a.XHILimit = t->XHILimit;
    a.lgCompton = ctx->lgCompton;
    a.colComp = 3 * t->nShells;
    a.colHit = 3 * t->nShells + 4;
    a.ncol = b.ncol;
    a.R = p_d;  // placeholder: W pointer kept in R until the launch (set per grid)
    b.set = true;
    return ensure_balance_grid(ctx);
}

int moc_update(moc_ctx* ctx, int ig, moc_update_stats* stats) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (g.attached && ig == 1 && ctx->dust.set && !ctx->dust.lgGas)  // updateCell without gas (:240-267)
        return moc_dust_temperature(ctx, ig, stats);
    if (!g.attached || ig != 1 || !ctx->balance.set || !g.d_ionDen) {
        ctx->err = "moc_update: needs an attached grid, moc_set_atomic and moc_set_balance";
        return MOC_ERR_STATE;
    }
    int rc = ensure_balance_grid(ctx);
    if (rc) return rc;
    MocBalance& b = ctx->balance;
    const double* W = nullptr;
    const MocBalanceArgs a = balance_args(ctx, &W);
    MOC_TRY(hipMemsetAsync(b.d_counters, 0, 4 * sizeof(unsigned long long), ctx->stream), "counters");
    // TeOld, NeOld = Te, Ne before the balance (iteration_mod.f90:1139-1142)
    const size_t C1b = ((size_t)g.nCells + 1) * sizeof(double);
    MOC_TRY(hipMemcpyAsync(g.d_TeOld, g.d_Te, C1b, hipMemcpyDeviceToDevice, ctx->stream), "TeOld");
    MOC_TRY(hipMemcpyAsync(g.d_NeOld, g.d_Ne, C1b, hipMemcpyDeviceToDevice, ctx->stream), "NeOld");
    MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    if ((rc = moc_launch_balance(ctx, a, g.d_J, W, b.ncol, g.d_rates, g.rates_chunk))) return rc;
    if (ctx->dust.set) {  // getDustT after iterateT (update_mod.f90:224-227)
        MocDust& d = ctx->dust;
        if ((rc = moc_launch_dust_temp(ctx, dust_args(ctx), g.d_J, d.d_W, d.nS * d.nA + 1, d.d_R)))
            return rc;
    }
    MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    MOC_TRY(hipEventSynchronize(ctx->ev1), "balance kernel");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    unsigned long long cnt[4];
    MOC_TRY(hipMemcpy(cnt, b.d_counters, sizeof(cnt), hipMemcpyDeviceToHost), "counters");
    g.f32_dirty = true;
    if (stats) {
        std::vector<int> conv((size_t)g.nCells + 1);
        MOC_TRY(hipMemcpy(conv.data(), g.d_conv, conv.size() * 4, hipMemcpyDeviceToHost), "conv");
        double nc = 0;
        for (int i = 1; i <= g.nCells; ++i) nc += conv[i];
        stats->noHit = (double)cnt[0];
        stats->noIonBal = (double)cnt[1];
        stats->noTeBal = (double)cnt[2];
        stats->converged = nc;
        stats->kernel_ms = ms;
    }
    return MOC_OK;
}

}  // extern "C"

namespace {

// updateCell arguments of the attached grid; *W receives the rate weights
MocBalanceArgs balance_args(moc_ctx* ctx, const double** W) {
    MocGrid& g = ctx->grid;
    MocBalance& b = ctx->balance;
    MocBalanceArgs a = b.a;
    *W = a.R;
    a.nCells = g.nCells;
    a.R = g.d_rates;
    a.Te = g.d_Te;
    a.Ne = g.d_Ne;
    a.Hden = g.d_Hden;
    a.ionDen = g.d_ionDen;
    a.lgConverged = g.d_conv;
    a.lgBlack = g.d_black;
    a.nIterT = g.d_nIterT;
    a.counters = b.d_counters;
    a.fitF32 = ctx->precision == MOC_PREC_MIXED ? 1 : 0;  // the precision split of the mixed form
    a.nGS = a.nGA = 0;
    if (ctx->dust.set && ctx->dust.lgPE) {  // grain terms inside the T iteration
        MocDust& d = ctx->dust;
        a.nGS = d.nS; a.nGA = d.nA; a.B = ctx->nbins;
        a.gAbun = d.d_abun; a.gWeight = d.d_weight; a.gRadius = d.d_radius; a.gVn = d.d_vn;
        a.gMsurf = d.d_msurf; a.gCabs = d.d_Cabs; a.gIp = d.d_ip; a.gNu = ctx->d_nu; a.J = g.d_J;
        a.Tdust = g.d_Tdust; a.Ndust = g.d_Ndust; a.collD = g.d_collD; a.peD = g.d_peD;
    }
    return a;
}

}  // namespace

extern "C" {

int moc_set_emission(moc_ctx* ctx, const moc_emission_tables* t) {
    if (!ctx) return MOC_ERR_ARG;
    drop_graph(ctx);
    if (!ctx->balance.set) { ctx->err = "moc_set_emission: call moc_set_balance first"; return MOC_ERR_STATE; }
    MocAtomic& at = ctx->atomic;
    const int B = ctx->nbins;
    if (!t || t->ntk < 2 || !t->tkGamma) { ctx->err = "moc_set_emission: invalid gamma tables"; return MOC_ERR_ARG; }
    for (int k = 0; k < 3; ++k) {
        if (t->nlim[k] < 2 || t->nlim[k] > MOC_EM_MAXLIM || (t->nlim[k] & 1) || !t->edgeP[k] || !t->logGamma[k]) {
            ctx->err = "moc_set_emission: gamma fit sizes (even, <= 64)";
            return MOC_ERR_ARG;
        }
        for (int i = 0; i < t->nlim[k]; ++i)
            if (t->edgeP[k][i] < 1 || t->edgeP[k][i] > B) { ctx->err = "moc_set_emission: edge pointer outside the mesh"; return MOC_ERR_ARG; }
        for (int p = 0; p < t->nlim[k] / 2; ++p)
            if (t->edgeP[k][2 * p] == t->edgeP[k][2 * p + 1]) {
                ctx->err = "moc_set_emission: frequency grid too coarse for the free-bound edges";
                return MOC_ERR_ARG;
            }
    }
    std::vector<int> slot(31, -1);
    for (int e = 0; e < at.nElemUsed; ++e) slot[at.elemZ[e]] = e;
    const bool heOn = slot[2] >= 0;
    if (t->HIPnuP < 1 || t->HIPnuP > B || !t->xsHI ||
        (heOn && (!t->xsHeI || !t->xsHeII || t->HeIPnuP < 1 || t->HeIPnuP > B || t->HeIIPnuP < 1 || t->HeIIPnuP > B))) {
        ctx->err = "moc_set_emission: H / He ground-state cross-sections";
        return MOC_ERR_ARG;
    }
    if (t->j2TsP < 1 || t->j2TsP > B || t->nFb < 0 || t->nFb > MOC_EM_MAXFB || t->nHlFiles < 1) {
        ctx->err = "moc_set_emission: invalid line pointers or table sizes";
        return MOC_ERR_ARG;
    }
    for (int k = 0; k < 4; ++k)
        if (t->lymP[k] < 0 || t->lymP[k] > B) { ctx->err = "moc_set_emission: He II Lyman bins"; return MOC_ERR_ARG; }
    for (int i = 0; i < 9 * 12; ++i)
        if (t->hlMap[i] < 0 || t->hlMap[i] >= t->nHlFiles) { ctx->err = "moc_set_emission: hlMap"; return MOC_ERR_ARG; }
    for (int i = 0; i < B; ++i)
        if (t->he2Row[i] < 0 || t->he2Row[i] > 40) { ctx->err = "moc_set_emission: he2Row"; return MOC_ERR_ARG; }
    std::vector<int> fbSlot((size_t)t->nFb);
    for (int n = 0; n < t->nFb; ++n) {
        const int z = t->fbZ[n], ion = t->fbIon[n];
        if (z < 3 || z > 19 || slot[z] < 0 || ion < 1 || ion >= at.nstages || t->fbLo[n] < 1 ||
            t->fbHi[n] > B || t->fbOff[n] < 0 ||
            (t->fbHi[n] > t->fbLo[n] && t->fbOff[n] + (t->fbHi[n] - t->fbLo[n]) > t->nFbXs)) {
            ctx->err = "moc_set_emission: heavy-ion free-bound entry";
            return MOC_ERR_ARG;
        }
        fbSlot[n] = slot[z];
    }
    // line emitters: hydrogenic ions (Z on, nstages > Z), n-level ions with data
    std::vector<int> hydZ, hydSlot, forSlot, forIon, forAtom, forTop;
    for (int z = 1; z <= 30; ++z)
        if (slot[z] >= 0 && at.nstages > z) { hydZ.push_back(z); hydSlot.push_back(slot[z]); }
    const std::vector<int>& aidx = ctx->balance.atomIdx;
    for (int z = 3; z <= 30; ++z) {
        if (slot[z] < 0) continue;
        const int top = z + 1 < at.nstages ? z + 1 : at.nstages;
        for (int ion = 1; ion <= top; ++ion) {
            const int atom = aidx[(size_t)z * 32 + ion];
            if (atom < 0) continue;
            forSlot.push_back(slot[z]); forIon.push_back(ion); forAtom.push_back(atom); forTop.push_back(top);
        }
    }
    free_emission(ctx->emission);
    MocEmission& em = ctx->emission;
    MocEmissionArgs& a = em.a;
    int rc;
    const double* pd;
#define EPUT(dst, src, n, what) do { if ((rc = em_put(ctx, src, (size_t)(n), &dst, what))) return rc; } while (0)
    EPUT(pd, ctx->widFlx.data(), B, "widFlx");
    a.widFlx = pd;
    EPUT(a.tk, t->tkGamma, t->ntk, "tkGamma");
    for (int k = 0; k < 3; ++k) {
        a.nlim[k] = t->nlim[k];
        EPUT(a.edge[k], t->edgeP[k], t->nlim[k], "edgeP");
        EPUT(a.lg[k], t->logGamma[k], (size_t)t->ntk * t->nlim[k], "logGamma");
    }
    {
        // the free-bound fit pair that covers each bin: the last pair p with
        // edge[2p-2] <= i < edge[2p-1] (fb_ff :437-507, a later pair overwrites)
        std::vector<int> gp((size_t)3 * B, 0);
        for (int k = 0; k < 3; ++k)
            for (int p = 1; p <= t->nlim[k] / 2; ++p)
                for (int i = t->edgeP[k][2 * p - 2]; i <= t->edgeP[k][2 * p - 1] - 1; ++i)
                    gp[(size_t)k * B + i - 1] = p;
        EPUT(a.gpair, gp.data(), gp.size(), "gpair");
    }
    EPUT(a.xsHI, t->xsHI, B - t->HIPnuP + 1, "xsHI");
    if (heOn) {
        EPUT(a.xsHeI, t->xsHeI, B - t->HeIPnuP + 1, "xsHeI");
        EPUT(a.xsHeII, t->xsHeII, B - t->HeIIPnuP + 1, "xsHeII");
    } else {
        a.xsHeI = nullptr;
        a.xsHeII = nullptr;
    }
    EPUT(a.fbSlot, fbSlot.data(), fbSlot.size(), "fbSlot");
    EPUT(a.fbIon, t->fbIon, t->nFb, "fbIon");
    EPUT(a.fbLo, t->fbLo, t->nFb, "fbLo");
    EPUT(a.fbHi, t->fbHi, t->nFb, "fbHi");
    EPUT(a.fbOff, t->fbOff, t->nFb, "fbOff");
    EPUT(a.fbXs, t->fbXs, t->nFbXs, "fbXs");
    EPUT(a.fbG, t->fbGratio, t->nFb, "fbGratio");
    EPUT(a.he2Row, t->he2Row, B, "he2Row");
    EPUT(a.hei2phot, t->hei2phot, 82, "hei2phot");
    EPUT(a.hydZ, hydZ.data(), hydZ.size(), "hydZ");
    EPUT(a.hydSlot, hydSlot.data(), hydSlot.size(), "hydSlot");
    EPUT(a.forSlot, forSlot.data(), forSlot.size(), "forSlot");
    EPUT(a.forIon, forIon.data(), forIon.size(), "forIon");
    EPUT(a.forAtom, forAtom.data(), forAtom.size(), "forAtom");
    EPUT(a.forTop, forTop.data(), forTop.size(), "forTop");
    {
        // ions of more than 5 levels (as cut at maxLev) are solved by teams of
        // 8, 16 or 32 lanes, in the level classes of nlevel_team_class
        std::vector<int> tl[kNlevTeams];
        const std::vector<int>& lev = ctx->balance.atomLevels;
        for (int k = 0; k < (int)forAtom.size(); ++k) {
            const int c = nlevel_team_class(lev[forAtom[k]]);
            if (c >= 0) tl[c].push_back(k);
        }
        for (int c = 0; c < kNlevTeams; ++c) {
            a.nForT[c] = (int)tl[c].size();
            EPUT(a.forT[c], tl[c].data(), tl[c].size(), "forT");
        }
    }
    EPUT(a.hlMap, t->hlMap, 9 * 12, "hlMap");
    EPUT(a.hlDens, t->hlDens, 14 * (size_t)t->nHlFiles, "hlDens");
    EPUT(a.hlLines, t->hlLines, 14 * 16 * 9 * (size_t)t->nHlFiles, "hlLines");
    EPUT(a.hbA, t->hbA, 36, "hbA");
    EPUT(a.hbB, t->hbB, 36, "hbB");
    EPUT(a.rbEdge, t->rbEdge, 2 * 3 * 26, "rbEdge");
    EPUT(a.r2aEdge, t->r2aEdge, 3 * 18, "r2aEdge");
    EPUT(a.heiRec, t->heiRec, 34 * 3 * 4, "heiRec");
    EPUT(a.heiiLyman, t->heiiLyman, 8, "heiiLyman");
#undef EPUT
    a.B = B;
    a.NE = at.nElemUsed;
    a.NS = at.nstages;
    a.nu = ctx->d_nu;
    a.log10nu = at.d_log10nu;
    a.abund = at.d_abund;
    a.ntk = t->ntk;
    a.HIP = t->HIPnuP;
    a.HeIP = heOn ? t->HeIPnuP : 0;
    a.HeIIP = heOn ? t->HeIIPnuP : 0;
    a.hSlot = slot[1];
    a.heSlot = slot[2];
    a.nFb = t->nFb;
    a.nHyd = (int)hydZ.size();
    a.nFor = (int)forAtom.size();
    a.j2TsP = t->j2TsP;
    for (int k = 0; k < 4; ++k) a.lymP[k] = t->lymP[k];
    em.set = true;
    return MOC_OK;
}

int moc_emission(moc_ctx* ctx, int ig, float* kernel_ms) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->emission.set || !ctx->balance.set || !g.d_ionDen) {
        ctx->err = "moc_emission: needs an attached grid, moc_set_balance and moc_set_emission";
        return MOC_ERR_STATE;
    }
    return run_emission(ctx, false, kernel_ms);
}

MocEmissionArgs emission_args(moc_ctx* ctx, bool skipUnchanged) {
    MocGrid& g = ctx->grid;
    MocEmissionArgs a = ctx->emission.a;
    a.bal = ctx->balance.a;
    a.nCells = g.nCells;
    a.C1 = g.nCells + 1;
    a.Te = g.d_Te;
    a.Ne = g.d_Ne;
    a.ionDen = g.d_ionDen;
    a.recPDF = g.d_recPDF;
    // the mixed form's continuum in FP32 (gas runs): the row goes to the FP32 plane
    a.recPDF32 = ctx->precision == MOC_PREC_MIXED && !ctx->dust.set ? g.f_recPDF : nullptr;
    a.totalLines = g.d_totalLines;
    a.lineSum = g.d_lineSum;
    a.TeOld = skipUnchanged ? g.d_TeOld : nullptr;
    a.NeOld = skipUnchanged ? g.d_NeOld : nullptr;
    a.nDS = a.nDA = 0;
    if (ctx->dust.set && ctx->dust.lgGas) {  // dust emission joins the diffuse CDF
        MocDust& d = ctx->dust;
        a.nDS = d.nS; a.nDA = d.nA;
        a.dAbun = d.d_abun; a.dWeight = d.d_weight; a.dTsub = d.d_tsub; a.dCabs = d.d_Cabs;
        a.Tdust = g.d_Tdust; a.Ndust = g.d_Ndust; a.Hden = g.d_Hden;
        a.qActive = d.lgQHeat && d.qActive && d.nQ > 0 ? 1 : 0;
        a.nQ = d.nQ; a.nTb = d.nTb; a.qIdx = d.d_qIdx; a.qT = g.d_qT; a.qP = g.d_qP;
    }
    return a;
}

int run_emission(moc_ctx* ctx, bool skipUnchanged, float* kernel_ms) {
    MocGrid& g = ctx->grid;
    const MocEmissionArgs a = emission_args(ctx, skipUnchanged);
    int rc;
    if (!a.recPDF32 && (rc = sync_f64_master(ctx, MOC_F_RECPDF))) return rc;
    if (ctx->dust.set && ctx->dust.lgGas && (rc = launch_qheat(ctx))) return rc;
    MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    rc = moc_launch_emission(ctx, a);
    if (rc) return rc;
    if (a.recPDF32) g.pdf32_master = true;
    MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    MOC_TRY(hipEventSynchronize(ctx->ev1), "emission kernel");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (kernel_ms) *kernel_ms = ms;
    g.f32_dirty = true;
    return MOC_OK;
}

}  // extern "C"

namespace {

void drop_graph(moc_ctx* ctx) {
    if (ctx->graphExec) (void)hipGraphExecDestroy(ctx->graphExec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graphExec = nullptr;
    ctx->graph = nullptr;
}

// The work of one iteration from the second on (cells whose Te and Ne moved
// by less than 0.5% keep their opacity and emission), enqueued on the
// context's stream with no waits, events, host copies or memsets, so that it
// can be captured as a graph: zero fills and device copies are kernels, the
// FP32 copies of the transport's inputs are refreshed unconditionally, and
// the transport reads its RNG counter base (nIterateMC * 256) from d_ctr.
// nBatch: the packet count of a batch of grids (views uploaded beforehand).
int enqueue_iteration(moc_ctx* ctx, uint64_t seed, int64_t nBatch) {
    MocGrid& g = ctx->grid;
    const bool dustOnly = ctx->dust.set && !ctx->dust.lgGas;
    const int64_t C1 = (int64_t)g.nCells + 1, n = C1 * ctx->nbins;
    int rc;
    if (dustOnly) {  // kappa and setDustPDF of every cell
        if ((rc = moc_launch_dust_opacity(ctx, dust_opacity_args(ctx))) ||
            (rc = moc_launch_dust_pdf(ctx, dust_args(ctx))))
            return rc;
    } else {
        const MocOpacityArgs oa = opacity_args(ctx, true);
        if (!oa.opacity32 && (rc = sync_f64_master(ctx, MOC_F_OPACITY))) return rc;
        if ((rc = moc_launch_opacity(ctx, oa))) return rc;
        if (oa.opacity32) g.op32_master = true;
        if (ctx->dust.set && (rc = moc_launch_dust_opacity(ctx, dust_opacity_args(ctx)))) return rc;
        const MocEmissionArgs ea = emission_args(ctx, true);
        if (!ea.recPDF32 && (rc = sync_f64_master(ctx, MOC_F_RECPDF))) return rc;
        if ((rc = moc_launch_emission(ctx, ea))) return rc;
        if (ea.recPDF32) g.pdf32_master = true;
    }
    // the tallies (moc_reset_tallies) and the transport counters
    moc_fill_zero_u64(reinterpret_cast<unsigned long long*>(g.d_J), n, ctx->stream);
    moc_fill_zero_u64(reinterpret_cast<unsigned long long*>(g.d_escaped), n, ctx->stream);
    moc_fill_zero_u64(ctx->d_counters, MOC_CNT_N, ctx->stream);
    moc_fill_zero_u64(reinterpret_cast<unsigned long long*>(ctx->d_Qphot), 1, ctx->stream);
    if (ctx->precision != MOC_PREC_FP64) {
        g.f32_dirty = true;
        if ((rc = refresh_f32(ctx))) return rc;
    }
    for (int s = 0; s < (g.batch() ? 1 : ctx->nStars); ++s) {
        const int64_t np = g.batch() ? nBatch : (int64_t)ctx->nPhotons[s];
        const double deltaE = ctx->Lstar[s] / ctx->nPhotons[s];  // photon_mod.f90:88
        if (!(deltaE >= 0.0)) { ctx->err = "moc_iterate: deltaEUsed < 0"; return MOC_ERR_ARG; }
        const uint32_t cb = (uint32_t)s + 1u;  // (nIterateMC * 256 from d_ctr) + iStar
        if (ctx->precision == MOC_PREC_FP64) {
            MocTransportArgs<double> a = transport_args_f64(ctx, s, np, seed, cb, deltaE);
            a.ctrDev = ctx->d_ctr;
            rc = moc_launch_transport_f64(ctx, a, nullptr);
        } else {
            MocTransportArgs<float> a = transport_args_f32(ctx, s, np, seed, cb, deltaE);
            a.ctrDev = ctx->d_ctr;
            rc = moc_launch_transport_f32(ctx, a, nullptr);
        }
        if (rc) return rc;
    }
    // moc_normalize_j
    moc_scale_f64(g.d_J, n, g.sym ? 1.e-9 / 8.0 : 1.e-9, ctx->stream);
    if (g.sym) moc_scale_f64(g.d_escaped, n, 1.0 / 8.0, ctx->stream);
    if (dustOnly) {  // moc_dust_temperature
        MocDust& d = ctx->dust;
        moc_fill_zero_u64(d.d_counters, 2, ctx->stream);
        return moc_launch_dust_temp(ctx, dust_args(ctx), g.d_J, d.d_W, d.nS * d.nA + 1, d.d_R);
    }
    // moc_update
    MocBalance& b = ctx->balance;
    const double* W = nullptr;
    const MocBalanceArgs a = balance_args(ctx, &W);
    moc_fill_zero_u64(b.d_counters, 4, ctx->stream);
    moc_copy_f64(g.d_Te, g.d_TeOld, C1, ctx->stream);
    moc_copy_f64(g.d_Ne, g.d_NeOld, C1, ctx->stream);
    if ((rc = moc_launch_balance(ctx, a, g.d_J, W, b.ncol, g.d_rates, g.rates_chunk))) return rc;
    if (ctx->dust.set) {
        MocDust& d = ctx->dust;
        if ((rc = moc_launch_dust_temp(ctx, dust_args(ctx), g.d_J, d.d_W, d.nS * d.nA + 1, d.d_R)))
            return rc;
    }
    MOC_TRY(hipGetLastError(), "iteration launch");
    return MOC_OK;
}

int capture_iteration(moc_ctx* ctx, uint64_t seed) {
    MocGrid& g = ctx->grid;
    const bool dustOnly = ctx->dust.set && !ctx->dust.lgGas;
    int rc;
    // what allocates or uploads from the host happens before the capture
    if (!ctx->d_ctr && (rc = dalloc_n(ctx, &ctx->d_ctr, 1, "iteration counter"))) return rc;
    if (!dustOnly && (rc = ensure_balance_grid(ctx))) return rc;
    if ((rc = ensure_hot(ctx))) return rc;
    int64_t nBatch = 0;
    if (g.batch() && (rc = upload_batch_views(ctx, 0, &nBatch))) return rc;
    if (g.nested && (rc = upload_nested_views(ctx, 0))) return rc;
    MOC_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    rc = enqueue_iteration(ctx, seed, nBatch);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    g.f32_dirty = true;  // the conversions were recorded, not run
    if (rc || e != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc) {
            ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
            rc = MOC_ERR_HIP;
        }
        return rc;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(graph);
        ctx->err = std::string("graph instantiate: ") + hipGetErrorString(e2);
        return MOC_ERR_HIP;
    }
    ctx->graph = graph;
    ctx->graphExec = exec;
    ctx->graphSeed = seed;
    ctx->graphOp32 = g.op32_master;
    ctx->graphPdf32 = g.pdf32_master;
    return MOC_OK;
}

// moc_iterate from the second iteration on, as one graph launch
int iterate_graph(moc_ctx* ctx, int nIterateMC, uint64_t seed, moc_iteration_stats* stats) {
    MocGrid& g = ctx->grid;
    const bool dustOnly = ctx->dust.set && !ctx->dust.lgGas;
    int rc;
    if (!ctx->graphExec || ctx->graphSeed != seed) {
        drop_graph(ctx);
        if ((rc = capture_iteration(ctx, seed))) return rc;
    }
    ctx->ctrHost = (uint32_t)nIterateMC * 256u;
    MOC_TRY(hipMemcpyAsync(ctx->d_ctr, &ctx->ctrHost, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream),
            "iteration counter");
    MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    MOC_TRY(hipGraphLaunch(ctx->graphExec, ctx->stream), "graph launch");
    // the replay wrote these planes in FP32 again
    g.op32_master = g.op32_master || ctx->graphOp32;
    g.pdf32_master = g.pdf32_master || ctx->graphPdf32;
    MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    MOC_TRY(hipEventSynchronize(ctx->ev1), "iteration graph");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    g.f32_dirty = true;
    unsigned long long cnt[MOC_CNT_N], ub[4] = {0, 0, 0, 0};
    MOC_TRY(hipMemcpy(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost), "counters");
    if (cnt[MOC_CNT_ERROR]) {
        ctx->err = "moc_transport: " + std::to_string(cnt[MOC_CNT_ERROR]) +
                   " packets reached an insane state (index outside the grid or emission "
                   "from an inactive cell)";
        return MOC_ERR_KERNEL;
    }
    if (!stats) return MOC_OK;
    double converged = 0.0;
    if (dustOnly) {
        MOC_TRY(hipMemcpy(ub, ctx->dust.d_counters, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost),
                "counters");
        converged = (double)ub[1];
        ub[1] = 0;
    } else {
        MOC_TRY(hipMemcpy(ub, ctx->balance.d_counters, sizeof(ub), hipMemcpyDeviceToHost), "counters");
        std::vector<int> conv((size_t)g.nCells + 1);
        MOC_TRY(hipMemcpy(conv.data(), g.d_conv, conv.size() * 4, hipMemcpyDeviceToHost), "conv");
        for (int i = 1; i <= g.nCells; ++i) converged += conv[i];
    }
    double packets = 0.0;
    for (int s = 0; s < ctx->nStars; ++s) packets += (double)(int64_t)ctx->nPhotons[s];
    const double nc = g.nCells > 0 ? (double)g.nCells : 1.0;
    stats->convPercent = 100.0 * converged / nc;
    stats->noHitPercent = 100.0 * (double)ub[0] / nc;
    stats->noIonBalPercent = 100.0 * (double)ub[1] / nc;
    stats->noTeBalPercent = 100.0 * (double)ub[2] / nc;
    stats->packets = packets;
    stats->opacity_ms = stats->emission_ms = stats->transport_ms = stats->update_ms = 0.0;
    stats->iteration_ms = ms;
    stats->segments = (double)cnt[MOC_CNT_SEGMENTS];
    stats->emissions = (double)cnt[MOC_CNT_EMISSIONS];
    stats->escaped = (double)cnt[MOC_CNT_ESCAPED];
    return MOC_OK;
}

}  // namespace

extern "C" {

int moc_set_graph(moc_ctx* ctx, int enable) {
    if (!ctx) return MOC_ERR_ARG;
    drop_graph(ctx);
    ctx->graphOn = enable != 0;
    return MOC_OK;
}

int moc_set_tuning(moc_ctx* ctx, const moc_tuning* t) {
    if (!ctx) return MOC_ERR_ARG;
    if (!t || t->emitBatch < 1 || t->emitBatch > 64 || t->segsPerTrip < 1 || t->segsPerTrip > 64 ||
        t->transportWaves < 64 || t->transportWaves > (1 << 20) || t->hotJ < 0 || t->hotJ > 1) {
        ctx->err = "moc_set_tuning: emitBatch and segsPerTrip in 1..64, transportWaves in 64..2^20, "
                   "hotJ 0 or 1";
        return MOC_ERR_ARG;
    }
    drop_graph(ctx);
    ctx->tune.emitBatch = t->emitBatch;
    ctx->tune.segsPerTrip = t->segsPerTrip;
    ctx->tune.transportWaves = t->transportWaves;
    ctx->tune.hotJ = t->hotJ;
    return MOC_OK;
}

int moc_get_tuning(moc_ctx* ctx, moc_tuning* t) {
    if (!ctx || !t) return MOC_ERR_ARG;
    t->emitBatch = ctx->tune.emitBatch;
    t->segsPerTrip = ctx->tune.segsPerTrip;
    t->transportWaves = ctx->tune.transportWaves;
    t->graph = ctx->graphOn ? 1 : 0;
    t->hotJ = ctx->tune.hotJ;
    return MOC_OK;
}

int moc_set_packets(moc_ctx* ctx, int iStar, int64_t n) {
    if (!ctx) return MOC_ERR_ARG;
    if (iStar < 1 || iStar > ctx->nStars || n < 1) { ctx->err = "moc_set_packets: iStar/n"; return MOC_ERR_ARG; }
    if (ctx->nPhotons[iStar - 1] != (double)n) drop_graph(ctx);
    ctx->nPhotons[iStar - 1] = (double)n;
    return MOC_OK;
}

int moc_iterate(moc_ctx* ctx, int ig, int nIterateMC, uint64_t seed, moc_iteration_stats* stats) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    const bool dustOnly = ctx->dust.set && !ctx->dust.lgGas;
    ctx->dust.firstIter = nIterateMC == 1;  // lgEquivalentTau: no dust opacity in iteration 1
    if (!g.attached || ig != 1 ||
        (!dustOnly && (!ctx->emission.set || !ctx->balance.set || !g.d_ionDen))) {
        ctx->err = "moc_iterate: needs an attached grid and either moc_set_atomic, moc_set_balance and "
                   "moc_set_emission, or moc_set_dust without gas";
        return MOC_ERR_STATE;
    }
    if (nIterateMC < 1 || nIterateMC > (1 << 23)) { ctx->err = "moc_iterate: nIterateMC"; return MOC_ERR_ARG; }
    // quantum heating from the second iteration on, once enough cells have
    // converged (setDiffusePDF :1701-1702, setDustPDF :1946-1947); those
    // iterations run eagerly (the switch is a host decision per iteration)
    const bool qheat = ctx->dust.set && ctx->dust.lgQHeat;
    if (qheat) ctx->dust.qActive = nIterateMC > 1 && ctx->lastConvPercent > ctx->dust.minConvQHeat;
    if (ctx->graphOn && nIterateMC > 1 && !qheat) return iterate_graph(ctx, nIterateMC, seed, stats);
    const bool skip = nIterateMC > 1;
    float tOp = 0.f, tEm = 0.f;
    int rc;
    if (dustOnly) {
        // dust kappa and setDustPDF of every cell (no skips without gas, :368-386)
        MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
        if ((rc = launch_dust_opacity(ctx))) return rc;
        MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
        MOC_TRY(hipEventSynchronize(ctx->ev1), "dust opacity");
        (void)hipEventElapsedTime(&tOp, ctx->ev0, ctx->ev1);
        MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
        if ((rc = launch_qheat(ctx)) || (rc = moc_launch_dust_pdf(ctx, dust_args(ctx)))) return rc;
        MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
        MOC_TRY(hipEventSynchronize(ctx->ev1), "dust emission");
        (void)hipEventElapsedTime(&tEm, ctx->ev0, ctx->ev1);
        g.f32_dirty = true;
    } else {
        if ((rc = run_opacity(ctx, skip, &tOp))) return rc;
        if ((rc = run_emission(ctx, skip, &tEm))) return rc;
    }
    if ((rc = moc_reset_tallies(ctx, ig))) return rc;
    double packets = 0.0, tTr = 0.0, seg = 0.0, emi = 0.0, esc = 0.0;
    if (g.batch()) {  // every grid's packets in one launch (RNG packet id (k << 32) + local)
        moc_transport_stats ts{};
        const uint32_t ctr = (uint32_t)nIterateMC * 256u + 1u;
        if ((rc = run_transport(ctx, 0, 0, seed, ctr, &ts))) return rc;
        for (int s = 0; s < ctx->nStars; ++s) packets += (double)(int64_t)ctx->nPhotons[s];
        tTr = ts.kernel_ms;
        seg = ts.segments;
        emi = ts.emissions;
        esc = ts.escaped;
    }
    for (int s = 1; !g.batch() && s <= ctx->nStars; ++s) {
        const int64_t n = (int64_t)ctx->nPhotons[s - 1];
        moc_transport_stats ts{};
        const uint32_t ctr = (uint32_t)nIterateMC * 256u + (uint32_t)s;
        if ((rc = moc_transport(ctx, s, n, seed, ctr, &ts))) return rc;
        packets += (double)n;
        tTr += ts.kernel_ms;
        seg += ts.segments;
        emi += ts.emissions;
        esc += ts.escaped;
    }
    if ((rc = moc_normalize_j(ctx, ig))) return rc;
    moc_update_stats us{};
    if ((rc = moc_update(ctx, ig, &us))) return rc;
    ctx->lastConvPercent = 100.0 * us.converged / (g.nCells > 0 ? (double)g.nCells : 1.0);
    if (stats) {
        const double nc = g.nCells > 0 ? (double)g.nCells : 1.0;
        stats->convPercent = 100.0 * us.converged / nc;
        stats->noHitPercent = 100.0 * us.noHit / nc;
        stats->noIonBalPercent = 100.0 * us.noIonBal / nc;
        stats->noTeBalPercent = 100.0 * us.noTeBal / nc;
        stats->packets = packets;
        stats->opacity_ms = tOp;
        stats->emission_ms = tEm;
        stats->transport_ms = tTr;
        stats->update_ms = us.kernel_ms;
        stats->iteration_ms = tOp + tEm + tTr + us.kernel_ms;
        stats->segments = seg;
        stats->emissions = emi;
        stats->escaped = esc;
    }
    return MOC_OK;
}

int moc_set_dust(moc_ctx* ctx, const moc_dust_tables* t) {
    if (!ctx) return MOC_ERR_ARG;
    drop_graph(ctx);
    if (ctx->grid.attached && ctx->grid.nGrids > 1) {
        ctx->err = "moc_set_dust: grains are not supported in a batch of grids or with sub-grids";
        return MOC_ERR_STATE;
    }
    MocGrid& g = ctx->grid;
    if (!g.attached) { ctx->err = "moc_set_dust: attach the grid first"; return MOC_ERR_STATE; }
    if (!t || t->nSpecies < 1 || t->nSizes < 1 || t->nTemps < 2 || !t->grainAbun || !t->grainWeight ||
        !t->TdustSublime || !t->Cabs || !t->Csca || !t->emIntegral || (t->lgGas != 0 && t->lgGas != 1)) {
        ctx->err = "moc_set_dust: invalid grain tables";
        return MOC_ERR_ARG;
    }
    if (t->lgQHeat && (!t->grainRadius || !t->grainKind || t->nTbins < 11 || t->nTbins > 2048 ||
                       !(t->Tmax > 1.) || !(t->minaQHeat >= 0.))) {
        ctx->err = "moc_set_dust: quantum heating needs the grain radii and kinds, 11..2048 "
                   "temperature bins and Tmax > 1 K";
        return MOC_ERR_ARG;
    }
    if (t->lgPhotoelectric && (!t->lgGas || !ctx->balance.set || !t->grainRadius || !t->grainVn ||
                               !t->MsurfAtom || !t->ipRyd)) {
        ctx->err = "moc_set_dust: the photoelectric terms need gas, moc_set_balance first and the "
                   "grain radii, work functions, surface masses and recombination energies";
        return MOC_ERR_ARG;
    }
    if ((int64_t)t->nSpecies * t->nSizes + 1 > 4096 ||
        (t->lgGas && (int64_t)t->nSpecies * t->nSizes > MOC_EM_MAXGRAIN)) {
        ctx->err = "moc_set_dust: too many (species, size) pairs (4095; 256 with gas)";
        return MOC_ERR_ARG;
    }
    free_dust(ctx->dust);
    MocDust& d = ctx->dust;
    const int B = ctx->nbins, nS = t->nSpecies, nA = t->nSizes, nT = t->nTemps, nG = nS * nA;
    const size_t C1 = (size_t)g.nCells + 1;
    d.nS = nS; d.nA = nA; d.nT = nT; d.lgGas = t->lgGas; d.XHILimit = t->XHILimit;
    int rc;
    // the absorption integrals sum_nu Cabs J/pi as rate columns, plus the hit test
    std::vector<double> W((size_t)B * (nG + 1));
    for (int b = 0; b < B; ++b) {
        for (int k = 0; k < nG; ++k) W[(size_t)b * (nG + 1) + k] = t->Cabs[(size_t)k * B + b] / 3.141592654;
        W[(size_t)b * (nG + 1) + nG] = 1.0;
    }
    if ((rc = dust_put(ctx, t->grainAbun, (size_t)nS, &d.d_abun, "grainAbun")) ||
        (rc = dust_put(ctx, t->grainWeight, (size_t)nA, &d.d_weight, "grainWeight")) ||
        (rc = dust_put(ctx, t->TdustSublime, (size_t)nS, &d.d_tsub, "TdustSublime")) ||
        (rc = dust_put(ctx, t->Cabs, (size_t)nG * B, &d.d_Cabs, "Cabs")) ||
        (rc = dust_put(ctx, t->Csca, (size_t)nG * B, &d.d_Csca, "Csca")) ||
        (rc = dust_put(ctx, t->emIntegral, (size_t)nG * nT, &d.d_emInt, "emIntegral")) ||
        (rc = dust_put(ctx, ctx->widFlx.data(), (size_t)B, &d.d_widFlx, "widFlx")) ||
        (rc = dust_put(ctx, W.data(), W.size(), &d.d_W, "dust W")) ||
        (rc = dust_put(ctx, (const double*)nullptr, (size_t)(nG + 1) * C1, &d.d_R, "dust R")) ||
        (rc = dust_put(ctx, (const unsigned long long*)nullptr, 2, &d.d_counters, "dust counters")))
        return rc;
    if (t->lgEquivalentTau) {
        const std::vector<double> zero((size_t)nS, 0.0);
        if ((rc = dust_put(ctx, zero.data(), (size_t)nS, &d.d_tsub0, "zero sublimation temperatures")))
            return rc;
    }
    d.lgQHeat = t->lgQHeat ? 1 : 0;
    d.equivTau = t->lgEquivalentTau ? 1 : 0;
    d.firstIter = false;
    if (d.lgQHeat) {
        // the grains below minaQHeat, species outer and sizes inner
        std::vector<int> qIdx((size_t)nG, -1), qG;
        for (int s = 0; s < nS; ++s)
            for (int ai = 0; ai < nA; ++ai)
                if (t->grainRadius[ai] < t->minaQHeat) {
                    qIdx[(size_t)s * nA + ai] = (int)qG.size();
                    qG.push_back(s * nA + ai);
                }
        d.nQ = (int)qG.size();
        d.nTb = t->nTbins;
        d.minaQHeat = t->minaQHeat;
        d.minConvQHeat = t->minConvQHeat;
        d.Tmax = t->Tmax;
        if ((rc = dust_put(ctx, qIdx.data(), qIdx.size(), &d.d_qIdx, "qHeat grains")) ||
            (rc = dust_put(ctx, qG.data(), qG.size(), &d.d_qG, "qHeat grains")) ||
            (rc = dust_put(ctx, t->grainKind, (size_t)nS, &d.d_kind, "grainKind")) ||
            (rc = dust_put(ctx, t->grainRadius, (size_t)nA, &d.d_radius, "grainRadius")))
            return rc;
        // the distributions of every cell, and the solver's work rows
        dfree(g.d_qT); dfree(g.d_qP); dfree(g.d_qWork);
        g.qBlocks = 0;
        if (d.nQ > 0) {
            const int blocks = kQHeatBlocks;
            if ((rc = dalloc_n(ctx, &g.d_qT, C1 * d.nQ * d.nTb, "qHeat T")) ||
                (rc = dalloc_n(ctx, &g.d_qP, C1 * d.nQ * d.nTb, "qHeat P")) ||
                (rc = dalloc_n(ctx, &g.d_qWork, (size_t)blocks * d.nTb * d.nTb, "qHeat work")))
                return rc;
            MOC_TRY(hipMemsetAsync(g.d_qT, 0, C1 * d.nQ * d.nTb * 8, ctx->stream), "zero");
            MOC_TRY(hipMemsetAsync(g.d_qP, 0, C1 * d.nQ * d.nTb * 8, ctx->stream), "zero");
            g.qBlocks = blocks;
        }
    }
    d.lgPE = t->lgPhotoelectric ? 1 : 0;
    if (d.lgPE &&
        ((!d.d_radius && (rc = dust_put(ctx, t->grainRadius, (size_t)nA, &d.d_radius, "grainRadius"))) ||
         (rc = dust_put(ctx, t->grainVn, (size_t)nS, &d.d_vn, "grainVn")) ||
         (rc = dust_put(ctx, t->MsurfAtom, (size_t)nS, &d.d_msurf, "MsurfAtom")) ||
         (rc = dust_put(ctx, t->ipRyd, (size_t)31 * 32, &d.d_ip, "ipRyd"))))
        return rc;
    // per-cell arrays of the grid: Ndust 0, Tdust 50 K (dust_mod.f90:66), kappa 0
    const size_t nTd = C1 * (nA + 1) * (nS + 1);
    dfree(g.d_Ndust); dfree(g.d_Tdust); dfree(g.d_absOpac); dfree(g.d_scaOpac);
    dfree(g.d_gasOpac); dfree(g.f_scaOpac); dfree(g.d_collD); dfree(g.d_peD);
    if (d.lgPE && ((rc = dalloc_n(ctx, &g.d_collD, C1 * nG, "collD")) ||
                   (rc = dalloc_n(ctx, &g.d_peD, C1 * nG, "peD"))))
        return rc;
    if (d.lgPE) {
        MOC_TRY(hipMemsetAsync(g.d_collD, 0, C1 * nG * 8, ctx->stream), "zero");
        MOC_TRY(hipMemsetAsync(g.d_peD, 0, C1 * nG * 8, ctx->stream), "zero");
    }
    if ((rc = dalloc_n(ctx, &g.d_Ndust, C1, "Ndust")) || (rc = dalloc_n(ctx, &g.d_Tdust, nTd, "Tdust")) ||
        (rc = dalloc_n(ctx, &g.d_absOpac, C1 * B, "absOpac")) ||
        (rc = dalloc_n(ctx, &g.d_scaOpac, C1 * B, "scaOpac")) ||
        (rc = dalloc_n(ctx, &g.f_scaOpac, C1 * B, "scaOpac32")) ||
        (t->lgGas && (rc = dalloc_n(ctx, &g.d_gasOpac, C1 * B, "gas opacity"))))
        return rc;
    if (!g.d_conv && ((rc = dalloc_n(ctx, &g.d_conv, C1, "lgConverged")) ||
                      (rc = dalloc_n(ctx, &g.d_black, C1, "lgBlack"))))
        return rc;
    std::vector<double> t50(nTd, 50.0);
    if ((rc = h2d(ctx, g.d_Tdust, t50.data(), nTd * 8, "Tdust"))) return rc;
    MOC_TRY(hipMemsetAsync(g.d_Ndust, 0, C1 * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_absOpac, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_scaOpac, 0, C1 * B * 8, ctx->stream), "zero");
    if (g.d_gasOpac) MOC_TRY(hipMemsetAsync(g.d_gasOpac, 0, C1 * B * 8, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_conv, 0, C1 * 4, ctx->stream), "zero");
    MOC_TRY(hipMemsetAsync(g.d_black, 0, C1 * 4, ctx->stream), "zero");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "moc_set_dust");
    d.set = true;
    g.f32_dirty = true;
    return MOC_OK;
}

int moc_dust_opacity(moc_ctx* ctx, int ig) {
    if (!ctx) return MOC_ERR_ARG;
    if (!ctx->grid.attached || ig != 1 || !ctx->dust.set) {
        ctx->err = "moc_dust_opacity: needs an attached grid and moc_set_dust";
        return MOC_ERR_STATE;
    }
    if (ctx->dust.lgGas) return run_opacity(ctx, false, nullptr);  // gas kappa, then the dust block
    int rc = launch_dust_opacity(ctx);
    if (rc) return rc;
    MOC_TRY(hipStreamSynchronize(ctx->stream), "dust opacity");
    return MOC_OK;
}

int moc_dust_emission(moc_ctx* ctx, int ig) {
    if (!ctx) return MOC_ERR_ARG;
    if (!ctx->grid.attached || ig != 1 || !ctx->dust.set || ctx->dust.lgGas) {
        ctx->err = "moc_dust_emission: needs an attached grid and moc_set_dust with lgGas = 0";
        return MOC_ERR_STATE;
    }
    int rc;
    if ((rc = launch_qheat(ctx)) || (rc = moc_launch_dust_pdf(ctx, dust_args(ctx)))) return rc;
    MOC_TRY(hipStreamSynchronize(ctx->stream), "dust emission");
    ctx->grid.f32_dirty = true;
    return MOC_OK;
}

int moc_set_qheat_state(moc_ctx* ctx, int nIterateMC, double convPercent) {
    if (!ctx) return MOC_ERR_ARG;
    MocDust& d = ctx->dust;
    if (!d.set || (!d.lgQHeat && !d.equivTau)) {
        ctx->err = "moc_set_qheat_state: moc_set_dust with lgQHeat or lgEquivalentTau first";
        return MOC_ERR_STATE;
    }
    d.qActive = nIterateMC > 1 && convPercent > d.minConvQHeat;
    d.firstIter = nIterateMC == 1;
    return MOC_OK;
}

int moc_dust_temperature(moc_ctx* ctx, int ig, moc_update_stats* stats) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->dust.set) {
        ctx->err = "moc_dust_temperature: needs an attached grid and moc_set_dust";
        return MOC_ERR_STATE;
    }
    MocDust& d = ctx->dust;
    MOC_TRY(hipMemsetAsync(d.d_counters, 0, 2 * sizeof(unsigned long long), ctx->stream), "counters");
    MOC_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    int rc = moc_launch_dust_temp(ctx, dust_args(ctx), g.d_J, d.d_W, d.nS * d.nA + 1, d.d_R);
    if (rc) return rc;
    MOC_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    MOC_TRY(hipEventSynchronize(ctx->ev1), "dust temperature");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    unsigned long long cnt[2];
    MOC_TRY(hipMemcpy(cnt, d.d_counters, sizeof(cnt), hipMemcpyDeviceToHost), "counters");
    if (stats) {
        stats->noHit = (double)cnt[0];
        stats->noIonBal = 0.;
        stats->noTeBal = 0.;
        stats->converged = (double)cnt[1];
        stats->kernel_ms = ms;
    }
    return MOC_OK;
}

}  // extern "C"

namespace {

// cell chunks of the volume-average partials: 4096 cells each, 1..256 of them
int volavg_chunks(int nCells) {
    const int n = (nCells + 4095) / 4096;
    return n < 1 ? 1 : (n > 256 ? 256 : n);
}

// Harrington et al. (1982) from the chunk-summed TW[2 * pair] = sum X Te Ne H dV,
// TW[2 * pair + 1] = sum X Ne H dV: <X> over the element's stages, <Te>
// weighted by the ion (output_mod.f90:682-730)
void volavg_finish(moc_ctx* ctx, const double* TW, double* TeVol, double* ionDenVol) {
    const int NE = ctx->atomic.nElemUsed, NS = ctx->atomic.nstages;
    for (int e = 0; e < NE; ++e) {
        const int top = std::min(ctx->atomic.elemZ[e] + 1, NS);
        double den = 0.0;
        for (int ion = 1; ion <= top; ++ion) den += TW[2 * ((size_t)(ion - 1) * NE + e) + 1];
        for (int ion = 1; ion <= NS; ++ion) {
            const size_t k = (size_t)(ion - 1) * NE + e;
            if (ion > top) { TeVol[k] = 0.0; ionDenVol[k] = 0.0; continue; }
            const double T = TW[2 * k], W = TW[2 * k + 1];
            TeVol[k] = W <= 0.0 ? 0.0 : T / W;
            ionDenVol[k] = den <= 0.0 ? 0.0 : W / den;
        }
    }
}

}  // namespace

extern "C" {

int moc_volume_averages(moc_ctx* ctx, int ig, double* TeVol, double* ionDenVol) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->atomic.set || !g.d_ionDen) {
        ctx->err = "moc_volume_averages: needs an attached grid and moc_set_atomic";
        return MOC_ERR_STATE;
    }
    if (g.batch()) {  // one nebula's output: sums over a batch would mix its grids
        ctx->err = "moc_volume_averages: the context holds a batch of grids; reduce one nebula per context";
        return MOC_ERR_STATE;
    }
    if (!TeVol || !ionDenVol) { ctx->err = "moc_volume_averages: null output"; return MOC_ERR_ARG; }
    const int NE = ctx->atomic.nElemUsed, NS = ctx->atomic.nstages, C1 = g.nCells + 1;
    const int nPairs = NE * NS;
    const int nChunks = volavg_chunks(g.nCells);
    double* d_part = nullptr;
    int rc;
    if ((rc = dalloc_n(ctx, &d_part, (size_t)nPairs * nChunks * 2, "volume-average partials"))) return rc;
    rc = moc_launch_volavg(ctx, g.d_ionDen, g.d_Te, g.d_Ne, g.d_Hden, g.d_vol, g.d_black, C1, nPairs,
                           nChunks, d_part);
    std::vector<double> part((size_t)nPairs * nChunks * 2);
    if (!rc) {
        const hipError_t e1 = hipMemcpyAsync(part.data(), d_part, part.size() * 8, hipMemcpyDeviceToHost,
                                             ctx->stream);
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            ctx->err = std::string("moc_volume_averages: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
            rc = MOC_ERR_HIP;
        }
    }
    dfree(d_part);
    if (rc) return rc;
    // sums over the chunks in order
    std::vector<double> TW((size_t)nPairs * 2, 0.0);
    for (int p = 0; p < nPairs; ++p)
        for (int c = 0; c < nChunks; ++c) {
            TW[2 * p] += part[((size_t)p * nChunks + c) * 2];
            TW[2 * p + 1] += part[((size_t)p * nChunks + c) * 2 + 1];
        }
    volavg_finish(ctx, TW.data(), TeVol, ionDenVol);
    return MOC_OK;
}

int moc_volume_averages_noalloc(moc_ctx* ctx, int ig, double* TeVol, double* ionDenVol) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->atomic.set || !g.d_ionDen) {
        ctx->err = "moc_volume_averages_noalloc: needs an attached grid and moc_set_atomic";
        return MOC_ERR_STATE;
    }
    if (g.batch()) {
        ctx->err = "moc_volume_averages_noalloc: the context holds a batch of grids; reduce one nebula per context";
        return MOC_ERR_STATE;
    }
    if (!TeVol || !ionDenVol) { ctx->err = "moc_volume_averages_noalloc: null output"; return MOC_ERR_ARG; }
    if (!ctx->d_scratch) { ctx->err = "moc_volume_averages_noalloc: no scratch buffer"; return MOC_ERR_STATE; }
    const int NE = ctx->atomic.nElemUsed, NS = ctx->atomic.nstages, C1 = g.nCells + 1;
    const int nPairs = NE * NS;
    const int nChunks = volavg_chunks(g.nCells);
    const size_t need = (size_t)nPairs * nChunks * 2 + (size_t)nPairs * 2;
    if (nPairs > MOC_MAXE * MOC_MAXS || need > kScratchDoubles) {
        ctx->err = "moc_volume_averages_noalloc: " + std::to_string(need) +
                   " doubles of partial sums exceed the fixed buffer of " + std::to_string(kScratchDoubles);
        return MOC_ERR_ARG;
    }
    double* d_part = ctx->d_scratch;
    double* d_sums = ctx->d_scratch + (size_t)nPairs * nChunks * 2;
    int rc = moc_launch_volavg(ctx, g.d_ionDen, g.d_Te, g.d_Ne, g.d_Hden, g.d_vol, g.d_black, C1, nPairs,
                               nChunks, d_part);
    if (!rc) rc = moc_launch_volavg_sum(ctx, d_part, nPairs, nChunks, C1, d_sums);
    if (rc) return rc;
    double TW[2 * MOC_MAXE * MOC_MAXS];
    const hipError_t e1 = hipMemcpyAsync(TW, d_sums, (size_t)nPairs * 2 * sizeof(double),
                                         hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        ctx->err = std::string("moc_volume_averages_noalloc: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
        return MOC_ERR_HIP;
    }
    volavg_finish(ctx, TW, TeVol, ionDenVol);
    return MOC_OK;
}

int moc_line_luminosities(moc_ctx* ctx, int ig, int* nIons, int* nLev, int* ionZ, int* ionStage,
                          double* lines, int64_t count, double* Hbeta) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1 || !ctx->emission.set || !ctx->balance.set || !g.d_ionDen) {
        ctx->err = "moc_line_luminosities: needs an attached grid, moc_set_balance and moc_set_emission";
        return MOC_ERR_STATE;
    }
    if (g.batch()) {  // one nebula's output: sums over a batch would mix its grids
        ctx->err = "moc_line_luminosities: the context holds a batch of grids; reduce one nebula per context";
        return MOC_ERR_STATE;
    }
    MocEmissionArgs a = ctx->emission.a;
    a.bal = ctx->balance.a;  // the n-level atoms
    const int L = a.bal.maxLev;
    if (L < 1) { ctx->err = "moc_line_luminosities: no n-level atoms (maxLev < 1)"; return MOC_ERR_STATE; }
    if (nIons) *nIons = a.nFor;
    if (nLev) *nLev = L;
    if (!lines) return MOC_OK;
    const int64_t nL = (int64_t)a.nFor * L * L;
    if (count < nL) { ctx->err = "moc_line_luminosities: output shorter than nIons*nLev*nLev"; return MOC_ERR_ARG; }
    if (ionZ || ionStage) {
        std::vector<int> slot(a.nFor), ion(a.nFor);
        if (a.nFor > 0) {
            MOC_TRY(hipMemcpy(slot.data(), a.forSlot, a.nFor * 4, hipMemcpyDeviceToHost), "ion list");
            MOC_TRY(hipMemcpy(ion.data(), a.forIon, a.nFor * 4, hipMemcpyDeviceToHost), "ion list");
        }
        for (int k = 0; k < a.nFor; ++k) {
            if (ionZ) ionZ[k] = ctx->atomic.elemZ[slot[k]];
            if (ionStage) ionStage[k] = ion[k];
        }
    }
    a.nCells = g.nCells;
    a.C1 = g.nCells + 1;
    a.Te = g.d_Te; a.Ne = g.d_Ne; a.ionDen = g.d_ionDen;
    int nChunks = (g.nCells + 127) / 128;
    nChunks = nChunks < 1 ? 1 : (nChunks > 1024 ? 1024 : nChunks);
    // ions in groups whose accumulators fit a workgroup's LDS; Hbeta with the first
    const int grp = moc_line_lum_group(a, L);
    const int nGroups = a.nFor > 0 ? (a.nFor + grp - 1) / grp : 1;
    const int64_t stride = (int64_t)(a.nFor < grp ? a.nFor : grp) * L * L + 1;
    double* d_part = nullptr;
    int rc;
    if ((rc = dalloc_n(ctx, &d_part, (size_t)nChunks * stride, "line partials"))) return rc;
    std::vector<double> part((size_t)nChunks * stride);
    for (int64_t k = 0; k < nL; ++k) lines[k] = 0.0;
    double hb = 0.0;
    for (int gi = 0; gi < nGroups && !rc; ++gi) {
        const int k0 = gi * grp;
        const int nk = a.nFor - k0 < grp ? a.nFor - k0 : grp;
        const int64_t n = (int64_t)nk * L * L;
        rc = moc_launch_line_lum(ctx, a, g.d_vol, g.d_black, g.d_Hden, nChunks, L, k0, nk, d_part);
        if (rc) break;
        const hipError_t e1 = hipMemcpyAsync(part.data(), d_part, (size_t)nChunks * (n + 1) * 8,
                                             hipMemcpyDeviceToHost, ctx->stream);
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            ctx->err = std::string("moc_line_luminosities: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
            rc = MOC_ERR_HIP;
            break;
        }
        double* out = lines + (int64_t)k0 * L * L;
        for (int c = 0; c < nChunks; ++c) {  // blocks in order: the sum does not depend on timing
            const double* pc = part.data() + (size_t)c * (n + 1);
            for (int64_t k = 0; k < n; ++k) out[k] += pc[k];
            if (gi == 0) hb += pc[n];
        }
    }
    dfree(d_part);
    if (rc) return rc;
    if (Hbeta) *Hbeta = hb;
    return MOC_OK;
}

int moc_sed(moc_ctx* ctx, int ig, double* sed, double* totalE) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_sed: no grid attached"; return MOC_ERR_STATE; }
    if (g.batch()) {  // one nebula's output: sums over a batch would mix its grids
        ctx->err = "moc_sed: the context holds a batch of grids; reduce one nebula per context";
        return MOC_ERR_STATE;
    }
    if (!sed) { ctx->err = "moc_sed: null output"; return MOC_ERR_ARG; }
    const int B = ctx->nbins;
    double* d_sed = nullptr;
    int rc;
    if ((rc = dalloc_n(ctx, &d_sed, (size_t)B, "SED"))) return rc;
    rc = moc_launch_sed(ctx, g.d_escaped, g.nCells + 1, B, d_sed);
    if (!rc) {
        const hipError_t e1 = hipMemcpyAsync(sed, d_sed, (size_t)B * 8, hipMemcpyDeviceToHost, ctx->stream);
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            ctx->err = std::string("moc_sed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
            rc = MOC_ERR_HIP;
        }
    }
    dfree(d_sed);
    if (rc) return rc;
    double tot = 0.0;
    for (int b = 0; b < B; ++b) {
        if (g.sym) sed[b] = sed[b] * 8.;  // the octant's tallies for the whole nebula (:3008)
        tot += sed[b];
    }
    if (totalE) *totalE = tot;
    return MOC_OK;
}

int moc_normalize_j(moc_ctx* ctx, int ig) {
    if (!ctx) return MOC_ERR_ARG;
    MocGrid& g = ctx->grid;
    if (!g.attached || ig != 1) { ctx->err = "moc_normalize_j: no grid attached"; return MOC_ERR_STATE; }
    const int64_t n = ((int64_t)g.nCells + 1) * ctx->nbins;
    // units: deltaE [1e36 erg/s] * l [cm] / V [1e45 cm^3] -> erg/s/cm^2 per bin
    moc_scale_f64(g.d_J, n, g.sym ? 1.e-9 / 8.0 : 1.e-9, ctx->stream);
    if (g.sym) moc_scale_f64(g.d_escaped, n, 1.0 / 8.0, ctx->stream);
    MOC_TRY(hipGetLastError(), "normalize");
    MOC_TRY(hipStreamSynchronize(ctx->stream), "normalize");
    return MOC_OK;
}

}  // extern "C"
