// C-ABI of the Hermite force + predictor back end (declared in
// include/hermite_hip.h): context, device buffers and kernel launches.
//
// Host-side flow per block step, mirroring the TITAN back end
// (SRC/hermite6-titan.cu:269-283, 495-538) on one HIP stream:
//   hh_set_jp             host write into the pinned j-store (no HIP call)
//   hh_predict_all        H2D of the slots written since the last upload,
//                         fused scatter + predict kernel over all slots
//   hh_force_first_nact   force kernel, D2H of nact Force records
#include "hh_kernels.hpp"
#include "hh_jstore.hpp"
#include "hh_plan_table.hpp"
#include "../../include/hermite_hip.h"

#include <algorithm>
#include <limits>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <cstddef>
#include <string>
#include <vector>

using namespace hh;

// One row of a force-launch tuning table: active counts up to nact_hi use
// this kernel (1 i-parallel, 2 j-parallel), i per lane and j-slice count.
struct TuneEntry {
    int nact_hi, kernel, ni, jsplit;
};
struct Geo {
    int kernel, ni, jsplit;
};

struct hh_ctx {
    int order = 0, nbody = 0, variant = 0, device = 0, npad = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *h_jp = nullptr;     // pinned, nbody records of jp_words
    double *h_force = nullptr;  // pinned, nbody Force records
    double *h_pot = nullptr;    // pinned, npad
    double *d_stage = nullptr;  // AoS upload target, nbody records
    double *d_jst = nullptr;    // SoA corrected j-store
    double *d_jpr = nullptr;    // SoA predicted j-particles
    double *d_force = nullptr;  // nbody Force records
    double *d_pot = nullptr;    // npad
    double *d_part = nullptr;   // partial sums of j-sliced launches
    size_t part_cap = 0;        // doubles
    // hh_force_first_nact_static: while set, j-sliced launches take their
    // partial sums from the device's static scratch (static_cap doubles),
    // with the j-slice count capped to fit, instead of growing d_part
    double *static_part = nullptr;
    size_t static_cap = 0;
    std::atomic<int> dirty_hi{0};
    int tune_kernel = 0, tune_split = 0, tune_ni = 0, tune_fold = 0;
    std::vector<TuneEntry> table;  // ascending nact_hi; empty = built-in choice
    int last_kernel = 0, last_split = 0, last_wgs = 0;
    std::string err;
};

static std::string g_create_err;

static int fail(hh_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

// The FP64 kernels' r^-1 (rsqrt64) element-wise, for checking it against a
// correctly rounded reference.
__global__ void rsqrt_eval_kernel(const double *__restrict__ x, double *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = hh::rsqrt64(x[i]);
}

// Scratch of hh_force_first_nact_static: the j-slice partial sums of one
// force call, in a static device array instead of a buffer that grows with
// hipMalloc.  Every GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name in
// host code is a stand-in, not a device address).
constexpr size_t HH_SCRATCH_DOUBLES = (size_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[HH_SCRATCH_DOUBLES];
constexpr int HH_MAX_DEVICES = 64;
static std::atomic<double *> g_scratch_ptr[HH_MAX_DEVICES];

#define HIPCHK(c, call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((c), HH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

static bool valid_order(int order) { return order == 4 || order == 6 || order == 8 || order == HH_ORDER_8DD; }
int hh_particle_words(int order) { return valid_order(order) ? p_words(order) : HH_E_ARG; }
int hh_force_words(int order) { return valid_order(order) ? f_words(order) : HH_E_ARG; }
int hh_jp_words(int order) { return valid_order(order) ? jp_words(order) : HH_E_ARG; }
int hh_pred_words(int order) { return valid_order(order) ? pr_words(order) : HH_E_ARG; }

const char *hh_last_error(const hh_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

}  // extern "C"

// ------------------------------------------------------------------ helpers
static void free_all(hh_ctx *c)
{
    // Teardown: errors are not actionable here, the buffers are released
    // either way.
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_stage);
    (void)hipFree(c->d_jst);
    (void)hipFree(c->d_jpr);
    (void)hipFree(c->d_force);
    (void)hipFree(c->d_pot);
    (void)hipFree(c->d_part);
    (void)hipHostFree(c->h_jp);
    (void)hipHostFree(c->h_force);
    (void)hipHostFree(c->h_pot);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

template <int ORDER>
static hipError_t launch_pad(hh_ctx *c)
{
    const int n = c->npad - c->nbody;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((pad_kernel<ORDER>), dim3((n + 255) / 256), dim3(256), 0, c->stream, c->nbody, c->npad,
                       c->d_jst, c->d_jpr);
    return hipGetLastError();
}

// Upload the slots written since the last upload and scatter them into the
// SoA store, predicting every slot to tsys when predict is true.
template <int ORDER, bool EX, bool PREDICT>
static hipError_t launch_predict(hh_ctx *c, int ndirty, double tsys)
{
    const int n = PREDICT ? c->npad : ndirty;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((predict_kernel<ORDER, EX, PREDICT>), dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->npad, ndirty, (const double *)c->d_stage, c->d_jst, c->d_jpr, tsys);
    return hipGetLastError();
}

static int upload_dirty(hh_ctx *c)
{
    const int nd = c->dirty_hi.exchange(0);
    if (nd > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_stage, c->h_jp, sizeof(double) * (size_t)nd * jp_words(c->order),
                                 hipMemcpyHostToDevice, c->stream));
    return nd;
}

template <bool PREDICT>
static int sync_and_predict(hh_ctx *c, double tsys)
{
    const int nd = upload_dirty(c);
    if (nd < 0) return nd;
    hipError_t e = hipSuccess;
    const bool ex = c->variant == HH_EXACT;
    switch (c->order) {
    case 4: e = ex ? launch_predict<4, true, PREDICT>(c, nd, tsys) : launch_predict<4, false, PREDICT>(c, nd, tsys); break;
    case 6: e = ex ? launch_predict<6, true, PREDICT>(c, nd, tsys) : launch_predict<6, false, PREDICT>(c, nd, tsys); break;
    case 8: e = ex ? launch_predict<8, true, PREDICT>(c, nd, tsys) : launch_predict<8, false, PREDICT>(c, nd, tsys); break;
    case 9: e = ex ? launch_predict<9, true, PREDICT>(c, nd, tsys) : launch_predict<9, false, PREDICT>(c, nd, tsys); break;
    }
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("predict launch: ") + hipGetErrorString(e));
    return HH_OK;
}

static int ensure_part(hh_ctx *c, size_t need)
{
    if (need <= c->part_cap) return HH_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(c->d_part));
    c->d_part = nullptr;
    c->part_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_part, need * sizeof(double)));
    c->part_cap = need;
    return HH_OK;
}

// Partial sums for a launch of nslice j-slices of `per_slice` doubles each:
// the context's growing buffer, or the static scratch during
// hh_force_first_nact_static, where the slice count is first capped to
// what the scratch holds (cap_slices).
static int cap_slices(const hh_ctx *c, int nslice, size_t per_slice)
{
    if (!c->static_part || nslice <= 1) return nslice;
    const size_t most = c->static_cap / per_slice;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nslice, most));
}

static int get_part(hh_ctx *c, size_t need, double *__restrict *part)
{
    if (c->static_part) {
        if (need > c->static_cap)
            return fail(c, HH_E_LIMIT, "partial sums of " + std::to_string(need) + " doubles exceed the static scratch of " +
                                          std::to_string(c->static_cap));
        *part = c->static_part;
        return HH_OK;
    }
    const int r = ensure_part(c, need);
    if (r) return r;
    *part = c->d_part;
    return HH_OK;
}

// Workgroups wanted per force launch: several per CU on 256 CUs.
static constexpr int TARGET_WG = 2048;

static hipError_t launch_reduce(hh_ctx *c, const double *part, int nslice, int nc, int f, int ipad, int ni,
                                double *fout, double *pout)
{
    const int per_block = 256 / RL;
    const int nt = ni * nc;
    hipLaunchKernelGGL(slice_reduce_kernel, dim3((nt + per_block - 1) / per_block), dim3(256), 0, c->stream, part,
                       nslice, nc, f, ipad, ni, fout, pout);
    return hipGetLastError();
}

// i-parallel launch of tile_kernel over src slots [0, ni) as i-particles.
template <int ORDER, int MODE, bool POT, int NI, int FOLD>
static int launch_tile_ni(hh_ctx *c, const Geo &g, const double *src, int ni, double eps2, double *fout, double *pout)
{
    constexpr int NC = n_acc(ORDER, POT);
    const int bi = NT * NI;
    const int iblocks = (ni + bi - 1) / bi;
    const int ntiles = (c->nbody + TJ - 1) / TJ;
    int nslice;
    if (MODE == EXACT) nslice = 1;
    else if (g.jsplit > 0) nslice = std::min(g.jsplit, ntiles);
    else nslice = std::max(1, std::min(ntiles, (TARGET_WG + iblocks - 1) / iblocks));
    nslice = cap_slices(c, nslice, (size_t)NC * iblocks * bi);
    const int tps = (ntiles + nslice - 1) / nslice;
    nslice = (ntiles + tps - 1) / tps;

    TileArgs a;
    a.src = src;
    a.npad = c->npad;
    a.nbody = c->nbody;
    a.ntiles = ntiles;
    a.tiles_per_slice = tps;
    a.nslice = nslice;
    a.nact = ni;
    a.ipad = iblocks * bi;
    a.eps2 = eps2;
    a.pot_mask_r2 = (ORDER == 0 && c->order == 4) ? 1 : 0;
    a.part = nullptr;
    a.fout = fout;
    a.pout = pout;
    if (nslice > 1) {
        const int r = get_part(c, (size_t)nslice * NC * a.ipad, &a.part);
        if (r) return r;
    }
    hipLaunchKernelGGL((tile_kernel<ORDER, MODE, NI, POT, FOLD>), dim3(iblocks, nslice), dim3(NT), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nslice > 1) e = launch_reduce(c, a.part, nslice, NC, ORDER == 0 ? 1 : f_words(ORDER), a.ipad, ni, fout, pout);
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("force launch: ") + hipGetErrorString(e));
    c->last_kernel = 1;
    c->last_split = nslice;
    c->last_wgs = iblocks * nslice;
    return HH_OK;
}

// Default FP32 partial-sum length of the SPLIT body.
static constexpr int DEFAULT_FOLD = 16;

template <int ORDER, int MODE, bool POT>
static int launch_tile(hh_ctx *c, const Geo &g, const double *src, int ni, double eps2, double *fout, double *pout)
{
    if constexpr (MODE == EXACT) {
        return launch_tile_ni<ORDER, MODE, POT, 1, 1>(c, g, src, ni, eps2, fout, pout);
    } else if constexpr (MODE == FP64) {
        if (g.ni == 2) return launch_tile_ni<ORDER, MODE, POT, 2, 1>(c, g, src, ni, eps2, fout, pout);
        return launch_tile_ni<ORDER, MODE, POT, 1, 1>(c, g, src, ni, eps2, fout, pout);
    } else {
        const int fold = c->tune_fold ? c->tune_fold : DEFAULT_FOLD;
        if (g.ni == 2) {
            switch (fold) {
            case 4: return launch_tile_ni<ORDER, MODE, POT, 2, 4>(c, g, src, ni, eps2, fout, pout);
            case 16: return launch_tile_ni<ORDER, MODE, POT, 2, 16>(c, g, src, ni, eps2, fout, pout);
            case 32: return launch_tile_ni<ORDER, MODE, POT, 2, 32>(c, g, src, ni, eps2, fout, pout);
            default: return launch_tile_ni<ORDER, MODE, POT, 2, 8>(c, g, src, ni, eps2, fout, pout);
            }
        }
        switch (fold) {
        case 4: return launch_tile_ni<ORDER, MODE, POT, 1, 4>(c, g, src, ni, eps2, fout, pout);
        case 16: return launch_tile_ni<ORDER, MODE, POT, 1, 16>(c, g, src, ni, eps2, fout, pout);
        case 32: return launch_tile_ni<ORDER, MODE, POT, 1, 32>(c, g, src, ni, eps2, fout, pout);
        default: return launch_tile_ni<ORDER, MODE, POT, 1, 8>(c, g, src, ni, eps2, fout, pout);
        }
    }
}

// Largest active count routed to the j-parallel kernel by default.
static constexpr int JPAR_MAX = 64;
static constexpr int TARGET_WG_JPAR = 1024;

// j-parallel launch (jpar_kernel) over [0, nact) as i-particles.
template <int ORDER, int MODE, int GI, bool POT>
static int launch_jpar(hh_ctx *c, const Geo &g, int nact, double eps2)
{
    constexpr int V = GI * n_acc(ORDER, POT);
    const int ngroups = (nact + GI - 1) / GI;
    const int ntj = c->npad / NT;
    int nslice = g.jsplit > 0 ? g.jsplit : (TARGET_WG_JPAR + ngroups - 1) / ngroups;
    nslice = std::max(1, std::min(nslice, ntj));
    nslice = cap_slices(c, nslice, (size_t)V * ngroups);
    const int jchunk = (ntj + nslice - 1) / nslice * NT;
    nslice = (c->npad + jchunk - 1) / jchunk;
    JParArgs a;
    a.src = c->d_jpr;
    a.npad = c->npad;
    a.nact = nact;
    a.ngroups = ngroups;
    a.jchunk = jchunk;
    a.nslice = nslice;
    a.eps2 = eps2;
    a.part = nullptr;
    a.fout = c->d_force;
    a.pout = c->d_pot;
    if (nslice > 1) {
        const int r = get_part(c, (size_t)nslice * V * ngroups, &a.part);
        if (r) return r;
    }
    hipLaunchKernelGGL((jpar_kernel<ORDER, MODE, GI, POT>), dim3(ngroups, nslice), dim3(NT), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nslice > 1)
        e = launch_reduce(c, a.part, nslice, n_acc(ORDER, POT), f_words(ORDER), ngroups * GI, nact, c->d_force,
                          c->d_pot);
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("force launch: ") + hipGetErrorString(e));
    c->last_kernel = 2;
    c->last_split = nslice;
    c->last_wgs = ngroups * nslice;
    return HH_OK;
}

template <int ORDER, int MODE, bool POT>
static int launch_jpar_gi(hh_ctx *c, const Geo &g, int nact, double eps2)
{
    if (nact == 1) return launch_jpar<ORDER, MODE, 1, POT>(c, g, nact, eps2);
    if (nact == 2) return launch_jpar<ORDER, MODE, 2, POT>(c, g, nact, eps2);
    return launch_jpar<ORDER, MODE, 4, POT>(c, g, nact, eps2);
}

// Launch geometry for an active count: the tuning table row covering nact
// (or the built-in choice), then any explicit hh_set_tuning overrides.
static Geo choose_geometry(const hh_ctx *c, int nact)
{
    Geo g{nact <= JPAR_MAX ? 2 : 1, 1, 0};
    for (const TuneEntry &t : c->table)
        if (nact <= t.nact_hi) {
            g = Geo{t.kernel, t.ni, t.jsplit};
            break;
        }
    if (c->tune_kernel) g.kernel = c->tune_kernel;
    if (c->tune_ni) g.ni = c->tune_ni;
    if (c->tune_split) g.jsplit = c->tune_split;
    return g;
}

template <int ORDER, int MODE, bool POT>
static int launch_force_mode(hh_ctx *c, int nact, double eps2)
{
    const Geo g = choose_geometry(c, nact);
    if constexpr (MODE != EXACT) {
        if (g.kernel == 2) return launch_jpar_gi<ORDER, MODE, POT>(c, g, nact, eps2);
    }
    return launch_tile<ORDER, MODE, POT>(c, g, c->d_jpr, nact, eps2, c->d_force, c->d_pot);
}

template <int ORDER>
static int launch_force_order(hh_ctx *c, int nact, double eps2, bool pot)
{
    switch (c->variant) {
    case HH_FP64:
        return pot ? launch_force_mode<ORDER, FP64, true>(c, nact, eps2) : launch_force_mode<ORDER, FP64, false>(c, nact, eps2);
    case HH_SPLIT:
        if constexpr (ORDER == 9) return fail(c, HH_E_ARG, "no split-precision kernel for double-double positions");
        else
        return pot ? launch_force_mode<ORDER, SPLIT, true>(c, nact, eps2) : launch_force_mode<ORDER, SPLIT, false>(c, nact, eps2);
    default:
        return pot ? launch_force_mode<ORDER, EXACT, true>(c, nact, eps2) : launch_force_mode<ORDER, EXACT, false>(c, nact, eps2);
    }
}

static int launch_force(hh_ctx *c, int nact, double eps2, bool pot)
{
    switch (c->order) {
    case 4: return launch_force_order<4>(c, nact, eps2, pot);
    case 6: return launch_force_order<6>(c, nact, eps2, pot);
    case 8: return launch_force_order<8>(c, nact, eps2, pot);
    default: return launch_force_order<9>(c, nact, eps2, pot);
    }
}

static int launch_potential(hh_ctx *c, double eps2)
{
    const Geo g{1, 1, 0};
    if (c->order == HH_ORDER_8DD)  // hi + lo positions, self excluded by index
        return c->variant == HH_EXACT ? launch_tile<1, EXACT, false>(c, g, c->d_jst, c->nbody, eps2, c->d_pot, nullptr)
                                      : launch_tile<1, FP64, false>(c, g, c->d_jst, c->nbody, eps2, c->d_pot, nullptr);
    // the potential is the energy diagnostic: FP64 for the split variant too,
    // so that its energy error measures the integration, not the diagnostic
    return c->variant == HH_EXACT ? launch_tile<0, EXACT, false>(c, g, c->d_jst, c->nbody, eps2, c->d_pot, nullptr)
                                  : launch_tile<0, FP64, false>(c, g, c->d_jst, c->nbody, eps2, c->d_pot, nullptr);
}

static int check_ctx(hh_ctx *c)
{
    if (!c) return HH_E_ARG;
    const hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return HH_OK;
}

// --------------------------------------------------------------- C-ABI
extern "C" {

int hh_create(int order, int nbody, int variant, int device, hh_ctx **out)
{
    if (!out) return fail(nullptr, HH_E_ARG, "out is null");
    *out = nullptr;
    if (!valid_order(order)) return fail(nullptr, HH_E_ARG, "order must be 4, 6, 8 or HH_ORDER_8DD");
    if (nbody <= 0) return fail(nullptr, HH_E_ARG, "nbody must be positive");
    if (variant < HH_FP64 || variant > HH_EXACT) return fail(nullptr, HH_E_ARG, "unknown variant");
    if (order == HH_ORDER_8DD && variant == HH_SPLIT)
        return fail(nullptr, HH_E_ARG, "no split-precision kernel for double-double positions");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(nullptr, HH_E_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, HH_E_ARG, "device ordinal out of range");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(nullptr, HH_E_NODEV, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, HH_E_NODEV, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, HH_E_NODEV, "hipSetDevice failed");

    hh_ctx *c = new hh_ctx();
    c->order = order;
    c->nbody = nbody;
    c->variant = variant;
    c->device = device;
    c->npad = (nbody + TJ - 1) / TJ * TJ;
    const size_t W = jp_words(order), Q = pr_words(order), F = f_words(order);
    const size_t np = (size_t)c->npad;
    auto bail = [&](const char *what, hipError_t e) {
        g_create_err = std::string(what) + ": " + hipGetErrorString(e);
        free_all(c);
        delete c;
        return HH_E_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail("event", e);
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail("event", e);
    if ((e = hipHostMalloc((void **)&c->h_jp, sizeof(double) * W * nbody)) != hipSuccess) return bail("host j-store", e);
    if ((e = hipHostMalloc((void **)&c->h_force, sizeof(double) * F * nbody)) != hipSuccess) return bail("host force", e);
    if ((e = hipHostMalloc((void **)&c->h_pot, sizeof(double) * np)) != hipSuccess) return bail("host pot", e);
    if ((e = hipMalloc(&c->d_stage, sizeof(double) * W * nbody)) != hipSuccess) return bail("stage", e);
    if ((e = hipMalloc(&c->d_jst, sizeof(double) * W * np)) != hipSuccess) return bail("j-store", e);
    if ((e = hipMalloc(&c->d_jpr, sizeof(double) * Q * np)) != hipSuccess) return bail("predictions", e);
    if ((e = hipMalloc(&c->d_force, sizeof(double) * F * nbody)) != hipSuccess) return bail("force", e);
    if ((e = hipMalloc(&c->d_pot, sizeof(double) * np)) != hipSuccess) return bail("pot", e);
    if ((e = hipMemsetAsync(c->d_jst, 0, sizeof(double) * W * np, c->stream)) != hipSuccess) return bail("memset", e);
    std::memset(c->h_jp, 0, sizeof(double) * W * nbody);
    switch (order) {
    case 4: e = launch_pad<4>(c); break;
    case 6: e = launch_pad<6>(c); break;
    case 8: e = launch_pad<8>(c); break;
    default: e = launch_pad<9>(c); break;
    }
    if (e != hipSuccess) return bail("pad", e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("sync", e);
    // the device's copy of the static scratch, looked up once per device
    // with that device current (it is)
    if (device < HH_MAX_DEVICES && !g_scratch_ptr[device].load()) {
        void *sp = nullptr;
        if ((e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch))) != hipSuccess) return bail("scratch address", e);
        g_scratch_ptr[device].store((double *)sp);
    }
    c->dirty_hi.store(nbody);
    *out = c;
    return HH_OK;
}

int hh_destroy(hh_ctx *c)
{
    if (!c) return HH_E_ARG;
    (void)hipSetDevice(c->device);
    free_all(c);
    delete c;
    return HH_OK;
}

double *hh_jstore(hh_ctx *c) { return c ? c->h_jp : nullptr; }

int hh_mark_dirty(hh_ctx *c, int lo, int hi)
{
    if (!c || lo < 0 || hi > c->nbody || lo > hi) return HH_E_ARG;
    dirty_raise(c->dirty_hi, hi);
    return HH_OK;
}

// set_jp: SRC/hermite6-titan.h:70-97 (hh_jstore.hpp).
int hh_set_jp(hh_ctx *c, int addr, const double *p)
{
    if (!c || !p || addr < 0 || addr >= c->nbody) return HH_E_ARG;
    jstore_write(c->order, c->h_jp, addr, p);
    dirty_raise(c->dirty_hi, addr + 1);
    return HH_OK;
}

int hh_predict_all(hh_ctx *c, double tsys)
{
    int r = check_ctx(c);
    if (r) return r;
    return sync_and_predict<true>(c, tsys);
}

int hh_force_first_nact(hh_ctx *c, int nact, double eps2, double *force, double *pot)
{
    int r = check_ctx(c);
    if (r) return r;
    if (nact < 0 || nact > c->nbody) return fail(c, HH_E_ARG, "nact out of range");
    if (nact > 0 && !force) return fail(c, HH_E_ARG, "force is null");
    if (nact == 0) return HH_OK;
    if ((r = launch_force(c, nact, eps2, pot != nullptr))) return r;
    const size_t F = f_words(c->order);
    HIPCHK(c, hipMemcpyAsync(c->h_force, c->d_force, sizeof(double) * F * nact, hipMemcpyDeviceToHost, c->stream));
    if (pot) HIPCHK(c, hipMemcpyAsync(c->h_pot, c->d_pot, sizeof(double) * nact, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(force, c->h_force, sizeof(double) * F * nact);
    if (pot) std::memcpy(pot, c->h_pot, sizeof(double) * nact);
    return HH_OK;
}

int hh_force_first_nact_static(hh_ctx *c, int nact, double eps2, double *force, double *pot)
{
    int r = check_ctx(c);
    if (r) return r;
    if (nact < 0 || nact > c->nbody) return fail(c, HH_E_ARG, "nact out of range");
    if (nact > 0 && !force) return fail(c, HH_E_ARG, "force is null");
    if (nact == 0) return HH_OK;
    double *sp = c->device < HH_MAX_DEVICES ? g_scratch_ptr[c->device].load() : nullptr;
    if (!sp) return fail(c, HH_E_STATE, "no static scratch address for this device");
    c->static_part = sp;
    c->static_cap = HH_SCRATCH_DOUBLES;
    r = launch_force(c, nact, eps2, pot != nullptr);
    c->static_part = nullptr;
    c->static_cap = 0;
    if (r) return r;
    const size_t F = f_words(c->order);
    HIPCHK(c, hipMemcpyAsync(c->h_force, c->d_force, sizeof(double) * F * nact, hipMemcpyDeviceToHost, c->stream));
    if (pot) HIPCHK(c, hipMemcpyAsync(c->h_pot, c->d_pot, sizeof(double) * nact, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(force, c->h_force, sizeof(double) * F * nact);
    if (pot) std::memcpy(pot, c->h_pot, sizeof(double) * nact);
    return HH_OK;
}

long long hh_static_scratch_doubles(void) { return (long long)HH_SCRATCH_DOUBLES; }

int hh_potential(hh_ctx *c, double eps2, double *pot)
{
    int r = check_ctx(c);
    if (r) return r;
    if (!pot) return fail(c, HH_E_ARG, "pot is null");
    if ((r = sync_and_predict<false>(c, 0.0))) return r;
    if ((r = launch_potential(c, eps2))) return r;
    HIPCHK(c, hipMemcpyAsync(c->h_pot, c->d_pot, sizeof(double) * c->nbody, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(pot, c->h_pot, sizeof(double) * c->nbody);
    return HH_OK;
}

int hh_set_tuning(hh_ctx *c, int kernel, int jsplit)
{
    if (!c || kernel < 0 || kernel > 2 || jsplit < 0) return HH_E_ARG;
    c->tune_kernel = kernel;
    c->tune_split = jsplit;
    return HH_OK;
}

int hh_set_tuning_table(hh_ctx *c, int n, const int *nact_hi, const int *kernel, const int *ni, const int *jsplit)
{
    if (!c || n < 0 || (n > 0 && (!nact_hi || !kernel || !ni || !jsplit))) return HH_E_ARG;
    std::vector<TuneEntry> t;
    for (int k = 0; k < n; k++) {
        if (kernel[k] < 1 || kernel[k] > 2 || ni[k] < 1 || ni[k] > 2 || jsplit[k] < 0) return HH_E_ARG;
        if (k > 0 && nact_hi[k] <= nact_hi[k - 1]) return HH_E_ARG;
        t.push_back(TuneEntry{nact_hi[k], kernel[k], ni[k], jsplit[k]});
    }
    c->table.swap(t);
    return HH_OK;
}

// Text table: one row per line, "nact_hi kernel ni jsplit", ascending.
int hh_load_tuning_file(hh_ctx *c, const char *path)
{
    if (!c || !path) return HH_E_ARG;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return fail(c, HH_E_ARG, std::string("cannot open tuning table ") + path);
    std::vector<int> a, b, d, e;
    char line[256];
    while (std::fgets(line, sizeof line, fp)) {
        int x0, x1, x2, x3;
        if (line[0] == '#') continue;
        if (std::sscanf(line, "%d %d %d %d", &x0, &x1, &x2, &x3) == 4) {
            a.push_back(x0);
            b.push_back(x1);
            d.push_back(x2);
            e.push_back(x3);
        }
    }
    std::fclose(fp);
    const int r = hh_set_tuning_table(c, (int)a.size(), a.data(), b.data(), d.data(), e.data());
    if (r) return fail(c, r, std::string("malformed tuning table ") + path);
    return HH_OK;
}

int hh_set_tuning_ex(hh_ctx *c, int i_per_thread, int split_fold)
{
    if (!c || i_per_thread < 0 || i_per_thread > 2) return HH_E_ARG;
    if (split_fold != 0 && split_fold != 4 && split_fold != 8 && split_fold != 16 && split_fold != 32) return HH_E_ARG;
    c->tune_ni = i_per_thread;
    c->tune_fold = split_fold;
    return HH_OK;
}

int hh_time_force(hh_ctx *c, int nact, double eps2, int reps, double *ms)
{
    int r = check_ctx(c);
    if (r) return r;
    if (nact <= 0 || nact > c->nbody || reps <= 0 || !ms) return fail(c, HH_E_ARG, "bad timing arguments");
    if ((r = launch_force(c, nact, eps2, false))) return r;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < reps; k++)
        if ((r = launch_force(c, nact, eps2, false))) return r;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / reps;
    return HH_OK;
}

int hh_time_predict(hh_ctx *c, double tsys, int reps, double *ms)
{
    int r = check_ctx(c);
    if (r) return r;
    if (reps <= 0 || !ms) return fail(c, HH_E_ARG, "bad timing arguments");
    if ((r = sync_and_predict<true>(c, tsys))) return r;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < reps; k++)
        if ((r = sync_and_predict<true>(c, tsys))) return r;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / reps;
    return HH_OK;
}

int hh_force_async(hh_ctx *c, int nact, double eps2, int with_pot)
{
    int r = check_ctx(c);
    if (r) return r;
    if (nact < 0 || nact > c->nbody) return fail(c, HH_E_ARG, "nact out of range");
    if (nact == 0) return HH_OK;
    return launch_force(c, nact, eps2, with_pot != 0);
}

int hh_predict_async(hh_ctx *c, double tsys)
{
    int r = check_ctx(c);
    if (r) return r;
    return sync_and_predict<true>(c, tsys);
}

int hh_synchronize(hh_ctx *c)
{
    int r = check_ctx(c);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HH_OK;
}

int hh_eval_rsqrt(hh_ctx *c, int n, const double *x, double *y)
{
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || (n > 0 && (!x || !y))) return fail(c, HH_E_ARG, "bad rsqrt arguments");
    if (n == 0) return HH_OK;
    double *d = nullptr;
    HIPCHK(c, hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
    const size_t bytes = sizeof(double) * (size_t)n;
    hipError_t e = hipMemcpyAsync(d, x, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        rsqrt_eval_kernel<<<(n + NT - 1) / NT, NT, 0, c->stream>>>(d, d + n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, d + n, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    const hipError_t ef = hipFree(d);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("hh_eval_rsqrt: ") + hipGetErrorString(e));
    return HH_OK;
}

int hh_last_geometry(const hh_ctx *c, int *kernel, int *jsplit, int *workgroups)
{
    if (!c) return HH_E_ARG;
    if (kernel) *kernel = c->last_kernel;
    if (jsplit) *jsplit = c->last_split;
    if (workgroups) *workgroups = c->last_wgs;
    return HH_OK;
}

}  // extern "C"

// =====================================================================
// GPU-resident block-step driver for one or many clusters (hhr_*).
// =====================================================================
#include "hh_resident.hpp"

#include <cmath>
#include <vector>

struct hhr_ctx {
    int order = 0, variant = 0, device = 0, nclu = 0, npad = 0, ntiles = 0, rw = 0;
    int force_grid = 0, reduce_grid = 0;
    // two-launch block steps for one cluster of at most fused_max
    // particles (FP64 and split): about fused_grid force workgroups,
    // fused_jpl j per lane (0: by size), the i-slice force kernel up to
    // fused_islice_max particles (0: by size, -1: never)
    int fused_max = R_FUSED_MAX, fused_grid = 512, fused_jpl = 0, fused_islice_max = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *rp = nullptr, *jpr = nullptr, *ipr = nullptr, *tn = nullptr, *part = nullptr, *pot = nullptr;
    int *ilist = nullptr, *tile_clu = nullptr;
    RClu *clu = nullptr;
    RGlob *glob = nullptr;
    size_t part_cap = 0;
    std::vector<int> off, n;
    std::vector<double> eps2;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    double last_run_ms = 0.0;
    long long steps_launched = 0;
    long long step_limit = 0;  // block steps per hhr_run call, 0 = none
    // Optional per-step timing: four events per step of a batch (step start,
    // before and after the force launch, step end).  Timed runs launch their
    // steps directly on the stream (no graph), so the events are ordinary
    // stream events.  The plan also writes each step's active count to
    // steplog, so every timed step is logged as (active count, force ms,
    // ms of the step's other launches).
    // Mode 2 times the same steps inside replayed graphs instead: the first
    // workgroup of every kernel writes the device clock at kernel start
    // (stamp, 4 words per step + the end of the batch), converted with the
    // device's wall-clock rate.
    int timing = 0;  // 0 off, 1 events around direct launches, 2 device clock stamps
    std::vector<hipEvent_t> kev;
    int *steplog = nullptr;
    int steplog_cap = 0;
    long long *stamp = nullptr;
    double wall_khz = 0.0;
    double force_ms = 0.0;
    long long force_launches = 0;
    std::vector<int> log_nact;
    std::vector<float> log_force, log_other;
    std::string err;
};

static std::string g_hhr_err;

static int rfail(hhr_ctx *c, int code, const std::string &m)
{
    if (c) c->err = m;
    else g_hhr_err = m;
    return code;
}

#define RCHK(c, call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return rfail((c), HH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static RArgs r_args(hhr_ctx *c)
{
    RArgs a;
    a.rp = c->rp;
    a.jpr = c->jpr;
    a.ipr = c->ipr;
    a.tn = c->tn;
    a.ilist = c->ilist;
    a.tile_clu = c->tile_clu;
    a.clu = c->clu;
    a.glob = c->glob;
    a.part = c->part;
    a.npad = c->npad;
    a.ntiles_total = c->ntiles;
    a.nclu = c->nclu;
    a.steplog = c->timing ? c->steplog : nullptr;
    a.steplog_cap = c->steplog_cap;
    return a;
}

static hipError_t r_mark(hhr_ctx *c, int k)
{
    return c->timing == 1 ? hipEventRecord(c->kev[k], c->stream) : hipSuccess;
}

// One block step: three launches (four for an ensemble); s is the step's
// index within its batch.  ST: the kernels' timed instances (timing mode 2).
template <int ORDER, int MODE, bool ST>
static hipError_t r_launch_step_t(hhr_ctx *c, int s)
{
    const RArgs a = r_args(c);
    hipError_t e = r_mark(c, 4 * s);
    if (e != hipSuccess) return e;
    // ensembles of up to R_PF_MAX clusters: planned inside the force launch
    const bool pf = c->nclu > 1 && c->nclu <= R_PF_MAX;
    if (c->nclu == 1 && c->ntiles <= R_PACK_MAX) {  // one cluster: planned by the compaction's last workgroup
        hipLaunchKernelGGL((r_compact_kernel<ORDER, MODE == EXACT, 1, ST>), dim3(c->ntiles), dim3(NT), 0, c->stream, a);
    } else if (c->nclu == 1) {
        hipLaunchKernelGGL((r_compact_kernel<ORDER, MODE == EXACT, 2, ST>), dim3(c->ntiles), dim3(NT), 0, c->stream, a);
    } else {
        hipLaunchKernelGGL((r_compact_kernel<ORDER, MODE == EXACT, 0, ST>), dim3(c->ntiles), dim3(NT), 0, c->stream, a);
        if (!pf) hipLaunchKernelGGL((r_plan_kernel<ORDER>), dim3(1), dim3(NT), 0, c->stream, a);
    }
    if ((e = r_mark(c, 4 * s + 1)) != hipSuccess) return e;
    if (pf)  // one more workgroup, for the step's bookkeeping
        hipLaunchKernelGGL((r_force_kernel<ORDER, MODE, ST, true>), dim3(c->force_grid + 1), dim3(NT), 0, c->stream, a);
    else hipLaunchKernelGGL((r_force_kernel<ORDER, MODE, ST, false>), dim3(c->force_grid), dim3(NT), 0, c->stream, a);
    if ((e = r_mark(c, 4 * s + 2)) != hipSuccess) return e;
    hipLaunchKernelGGL((r_reduce_correct_kernel<ORDER, ST>), dim3(c->reduce_grid), dim3(NT), 0, c->stream, a);
    if ((e = r_mark(c, 4 * s + 3)) != hipSuccess) return e;
    return hipGetLastError();
}

template <int ORDER, int MODE>
static hipError_t r_launch_step(hhr_ctx *c, int s)
{
    return c->timing == 2 ? r_launch_step_t<ORDER, MODE, true>(c, s) : r_launch_step_t<ORDER, MODE, false>(c, s);
}

// One block step of a small single cluster in two launches (list + force,
// correct + bookkeeping); the events bracket the first as the force launch.
// Its partials (one per j-slice and slot) fit in the partial buffer, which
// holds at least NT * R_TASK_MAX force records.
static_assert((size_t)R_FUSED_SLICES * R_FUSED_MAX <= (size_t)NT * R_TASK_MAX, "fused partials");
template <int ORDER, int MODE, int JPL, bool ST>
static hipError_t r_launch_fused_jpl(hhr_ctx *c, int s)
{
    const RArgs a = r_args(c);
    const int n = c->n[0], off = c->off[0];
    const int S = (n + JPL * NT - 1) / (JPL * NT);   // j-slices
    const int wps = std::max(1, c->fused_grid / S);  // workgroups per slice
    hipError_t e;
    if ((e = r_mark(c, 4 * s)) != hipSuccess || (e = r_mark(c, 4 * s + 1)) != hipSuccess) return e;
    // i-slice force kernel (no dependent round trip) up to 1024 particles,
    // 2048 at 4th order; above, the active list built in every workgroup
    // balances crowded steps better (measured on MI355X, orders 4/6/8)
    const bool islice = c->fused_islice_max ? n <= c->fused_islice_max : (n <= 1024 || (ORDER == 4 && n <= 2048));
    if (islice) {
        const int Si = (n + NT - 1) / NT;
        const int wpc = std::max(1, c->fused_grid / (Si * S));  // workgroups per (i-slice, j-slice)
        hipLaunchKernelGGL((r_fused_islice_kernel<ORDER, MODE, JPL, ST>), dim3(wpc * Si * S), dim3(NT), 0, c->stream, a, n, off, Si, S,
                           wpc);
    } else {
        hipLaunchKernelGGL((r_fused_force_kernel<ORDER, MODE, JPL, ST>), dim3(wps * S), dim3(NT), 0, c->stream, a, n, off, S, wps);
    }
    if ((e = r_mark(c, 4 * s + 2)) != hipSuccess) return e;
    hipLaunchKernelGGL((r_fused_correct_kernel<ORDER, ST>), dim3((n + NT - 1) / NT), dim3(NT), 0, c->stream, a, n, off, S);
    if ((e = r_mark(c, 4 * s + 3)) != hipSuccess) return e;
    return hipGetLastError();
}

// j per lane of the force launch; 0 picks by size: two up to 1024
// particles and at 8th order (four would leave one wave per SIMD), four
// above (measured on MI355X at N = 1024, 2048 and 4096)
template <int ORDER, int MODE>
static hipError_t r_launch_fused_step(hhr_ctx *c, int s)
{
    const int n = c->n[0];
    const int jpl = c->fused_jpl ? c->fused_jpl : (n <= 1024 || ORDER == 8) ? 2 : 4;
    if (c->timing == 2) {
        if (jpl == 1) return r_launch_fused_jpl<ORDER, MODE, 1, true>(c, s);
        if (jpl == 2) return r_launch_fused_jpl<ORDER, MODE, 2, true>(c, s);
        return r_launch_fused_jpl<ORDER, MODE, 4, true>(c, s);
    }
    if (jpl == 1) return r_launch_fused_jpl<ORDER, MODE, 1, false>(c, s);
    if (jpl == 2) return r_launch_fused_jpl<ORDER, MODE, 2, false>(c, s);
    return r_launch_fused_jpl<ORDER, MODE, 4, false>(c, s);
}

static bool r_fused(const hhr_ctx *c)
{
    return c->nclu == 1 && c->variant != HH_EXACT && c->n[0] <= c->fused_max;
}

static hipError_t r_step(hhr_ctx *c, int s)
{
    if (r_fused(c)) {
        switch (c->order * 4 + c->variant) {
        case 16 + HH_FP64: return r_launch_fused_step<4, FP64>(c, s);
        case 16 + HH_SPLIT: return r_launch_fused_step<4, SPLIT>(c, s);
        case 24 + HH_FP64: return r_launch_fused_step<6, FP64>(c, s);
        case 24 + HH_SPLIT: return r_launch_fused_step<6, SPLIT>(c, s);
        case 32 + HH_FP64: return r_launch_fused_step<8, FP64>(c, s);
        default: return r_launch_fused_step<8, SPLIT>(c, s);
        }
    }
    switch (c->order * 4 + c->variant) {
    case 16 + HH_FP64: return r_launch_step<4, FP64>(c, s);
    case 16 + HH_SPLIT: return r_launch_step<4, SPLIT>(c, s);
    case 16 + HH_EXACT: return r_launch_step<4, EXACT>(c, s);
    case 24 + HH_FP64: return r_launch_step<6, FP64>(c, s);
    case 24 + HH_SPLIT: return r_launch_step<6, SPLIT>(c, s);
    case 24 + HH_EXACT: return r_launch_step<6, EXACT>(c, s);
    case 32 + HH_FP64: return r_launch_step<8, FP64>(c, s);
    case 32 + HH_SPLIT: return r_launch_step<8, SPLIT>(c, s);
    default: return r_launch_step<8, EXACT>(c, s);
    }
}

static void r_free_events(hhr_ctx *c)
{
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    c->kev.clear();
    (void)hipFree(c->steplog);
    c->steplog = nullptr;
    c->steplog_cap = 0;
    (void)hipFree(c->stamp);
    c->stamp = nullptr;
}

// A graph captured with other kernel arguments (timing buffers) is stale.
static void r_drop_graph(hhr_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_steps = 0;
}

// The clock at the end of a batch (timing mode 2): its last step's end.
static hipError_t r_stamp_end(hhr_ctx *c)
{
    if (c->timing != 2) return hipSuccess;
    hipLaunchKernelGGL(r_stamp_end_kernel, dim3(1), dim3(1), 0, c->stream, r_args(c));
    return hipGetLastError();
}

static void r_free(hhr_ctx *c)
{
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    (void)hipFree(c->rp);
    (void)hipFree(c->jpr);
    (void)hipFree(c->ipr);
    (void)hipFree(c->tn);
    (void)hipFree(c->part);
    (void)hipFree(c->pot);
    (void)hipFree(c->ilist);
    (void)hipFree(c->tile_clu);
    (void)hipFree(c->clu);
    (void)hipFree(c->glob);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    r_free_events(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// The measured plan of this order and kind (one cluster or an ensemble)
// nearest in total particles, within a factor of 4 (hh_plan_table.hpp);
// the built-in plan otherwise.
static RPlan measured_plan(int order, int nclu, long long n_total)
{
    RPlan p;
    double best = 2.0 + 1e-9;
    for (const PlanRow &r : kPlanTable) {
        if (r.order != order || (r.nclu > 1) != (nclu > 1)) continue;
        const double d = std::fabs(std::log2((double)r.n_total) - std::log2((double)n_total));
        if (d <= best) {
            best = d;
            p.jpar_max = r.v[0];
            p.smax = r.v[1];
            p.task_target = r.v[2];
            p.jpar_target = r.v[3];
            p.wj4_max = r.v[4];
            p.wj2_max = r.v[5];
        }
    }
    return p;
}

extern "C" {

const char *hhr_last_error(const hhr_ctx *c) { return c ? c->err.c_str() : g_hhr_err.c_str(); }

// recs: the clusters' reference Particle records, concatenated, after
// init_step (stored derivatives and dt set).  params: per cluster
// {eps2, eta, dtmax, dtmin, tsys}.
int hhr_create(int order, int variant, int device, int nclu, const int *ns, const double *recs,
               const double *params, hhr_ctx **out)
{
    if (!out || !ns || !recs || !params || nclu <= 0) return rfail(nullptr, HH_E_ARG, "null argument");
    *out = nullptr;
    if (order != 4 && order != 6 && order != 8) return rfail(nullptr, HH_E_ARG, "order must be 4, 6 or 8");
    if (variant < HH_FP64 || variant > HH_EXACT) return rfail(nullptr, HH_E_ARG, "unknown variant");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return rfail(nullptr, HH_E_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return rfail(nullptr, HH_E_ARG, "device ordinal out of range");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return rfail(nullptr, HH_E_NODEV, "device is not gfx950");
    if (hipSetDevice(device) != hipSuccess) return rfail(nullptr, HH_E_NODEV, "hipSetDevice failed");

    hhr_ctx *c = new hhr_ctx();
    c->order = order;
    c->variant = variant;
    c->device = device;
    c->nclu = nclu;
    c->rw = r_words(order);
    c->off.resize(nclu);
    c->n.assign(ns, ns + nclu);
    c->eps2.resize(nclu);
    int tot = 0;
    for (int k = 0; k < nclu; k++) {
        if (ns[k] <= 0) {
            delete c;
            return rfail(nullptr, HH_E_ARG, "cluster sizes must be positive");
        }
        c->off[k] = tot;
        tot += (ns[k] + TJ - 1) / TJ * TJ;
    }
    c->npad = tot;
    c->ntiles = tot / TJ;
    const int pw = p_words(order), rw = c->rw, Q = pr_words(order), F = f_words(order);

    // host SoA image of the resident records
    std::vector<double> img((size_t)rw * tot, 0.0), tn(tot, std::numeric_limits<double>::infinity());
    std::vector<int> tclu(c->ntiles);
    std::vector<RClu> hc(nclu);
    size_t src = 0;
    for (int k = 0; k < nclu; k++) {
        const int o = c->off[k], nk = ns[k];
        const int ntk = (nk + TJ - 1) / TJ;
        for (int t = 0; t < ntk; t++) tclu[o / TJ + t] = k;
        for (int i = 0; i < (ntk * TJ); i++) {
            const size_t s = (size_t)o + i;
            if (i >= nk) {  // padding: massless, far away, never active
                for (int w = 0; w < 3; w++) img[(size_t)w * tot + s] = FAR;
                continue;
            }
            const double *p = recs + (src + i) * pw;
            auto put = [&](int w, double v) { img[(size_t)w * tot + s] = v; };
            for (int d = 0; d < 3; d++) {
                put(0 + d, p[4 + d]);    // pos
                put(4 + d, p[7 + d]);    // vel
                put(8 + d, p[10 + d]);   // acc
                put(11 + d, p[13 + d]);  // jrk
                put(14 + d, p[16 + d]);  // snp
                put(17 + d, p[19 + d]);  // crk
                if (order >= 6) {
                    put(20 + d, p[22 + d]);  // d4a
                    put(23 + d, p[25 + d]);  // d5a
                }
                if (order == 8) {
                    put(R_D6A + d, p[28 + d]);
                    put(R_D7A + d, p[31 + d]);
                }
            }
            put(3, p[1]);     // mass
            put(7, p[2]);     // tlast
            put(R_DT, p[3]);  // dt
            put(R_ID, p[0]);  // id bits
            tn[s] = p[2] + p[3];
        }
        src += nk;
        const double *pr = params + 5 * k;
        RClu &h = hc[k];
        std::memset(&h, 0, sizeof(h));
        h.eps2 = pr[0];
        h.etapow = std::pow(pr[1], double(2 * (order - 3)));
        h.dtmax = pr[2];
        h.dtmin = pr[3];
        h.tsys = pr[4];
        h.tend = pr[4];
        h.tnext = pr[4];
        // first block time; the minima are atomicMin on the bit patterns,
        // which order like the values only for positive doubles
        double tmin = INFINITY;
        for (int i = 0; i < nk; i++) tmin = std::min(tmin, tn[(size_t)o + i]);
        if (!(tmin > 0.0) || !std::isfinite(tmin) || pr[4] < 0.0) {
            delete c;
            return rfail(nullptr, HH_E_ARG, "block times tlast + dt must be positive and finite, tsys >= 0");
        }
        for (int q = 0; q < R_SLOTS; q++) h.tmin_slot[0][q] = h.tmin_slot[1][q] = 0x7ff0000000000000ULL;
        std::memcpy(&h.tmin_slot[0][0], &tmin, sizeof(double));
        h.off = o;
        h.n = nk;
        h.ntiles = ntk;
        c->eps2[k] = pr[0];
    }
    RGlob g;
    std::memset(&g, 0, sizeof(g));
    g.nclu = nclu;
    g.exact = variant == HH_EXACT;
    g.plan = measured_plan(order, nclu, (long long)src);  // src: the clusters' particles in total

    auto bail = [&](const char *what, hipError_t e) {
        g_hhr_err = std::string(what) + ": " + hipGetErrorString(e);
        r_free(c);
        delete c;
        return HH_E_HIP;
    };
    hipError_t e;
    // Partials of one block step: i-parallel clusters hold nslice * ngrp <=
    // budget + ngrp tasks of G = NT * r_ni i each, budgets summing to at
    // most task_target + nclu, and ngrp summing to at most (tot + G nclu) / G;
    // j-parallel clusters need less.  Sized for the largest task target.
    const size_t G = (size_t)NT * r_ni(FP64);
    c->part_cap = (size_t)F * (tot + G * nclu) + (size_t)F * G * (R_TASK_MAX + nclu);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail("event", e);
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail("event", e);
    if ((e = hipMalloc(&c->rp, sizeof(double) * img.size())) != hipSuccess) return bail("records", e);
    if ((e = hipMalloc(&c->jpr, sizeof(double) * (size_t)Q * tot)) != hipSuccess) return bail("predictions", e);
    if ((e = hipMalloc(&c->ipr, sizeof(double) * (size_t)Q * tot)) != hipSuccess) return bail("active predictions", e);
    if ((e = hipMalloc(&c->tn, sizeof(double) * tot)) != hipSuccess) return bail("tn", e);
    if ((e = hipMalloc(&c->pot, sizeof(double) * tot)) != hipSuccess) return bail("pot", e);
    if ((e = hipMalloc(&c->part, sizeof(double) * c->part_cap)) != hipSuccess) return bail("partials", e);
    if ((e = hipMalloc(&c->ilist, sizeof(int) * tot)) != hipSuccess) return bail("ilist", e);
    if ((e = hipMalloc(&c->tile_clu, sizeof(int) * c->ntiles)) != hipSuccess) return bail("tile map", e);
    if ((e = hipMalloc(&c->clu, sizeof(RClu) * nclu)) != hipSuccess) return bail("clusters", e);
    if ((e = hipMalloc(&c->glob, sizeof(RGlob))) != hipSuccess) return bail("glob", e);
    if ((e = hipMemcpy(c->rp, img.data(), sizeof(double) * img.size(), hipMemcpyHostToDevice)) != hipSuccess) return bail("upload", e);
    if ((e = hipMemcpy(c->tn, tn.data(), sizeof(double) * tot, hipMemcpyHostToDevice)) != hipSuccess) return bail("upload", e);
    if ((e = hipMemcpy(c->tile_clu, tclu.data(), sizeof(int) * c->ntiles, hipMemcpyHostToDevice)) != hipSuccess) return bail("upload", e);
    if ((e = hipMemcpy(c->clu, hc.data(), sizeof(RClu) * nclu, hipMemcpyHostToDevice)) != hipSuccess) return bail("upload", e);
    if ((e = hipMemcpy(c->glob, &g, sizeof(RGlob), hipMemcpyHostToDevice)) != hipSuccess) return bail("upload", e);
    if ((e = hipMemset(c->jpr, 0, sizeof(double) * (size_t)Q * tot)) != hipSuccess) return bail("memset", e);
    // sixteen workgroups per CU, several times those resident at once, so
    // that the dispatcher balances the work as workgroups finish: a step's
    // force tasks (up to the 8192 of the plan's target) mostly get a
    // workgroup each.  Measured faster than 4 per CU (C2 1-2 %, C4 4 %) and
    // no slower than 32
    {
        int khz = 0;
        if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device)) != hipSuccess)
            return bail("wall clock rate", e);
        c->wall_khz = khz;
    }
    c->force_grid = prop.multiProcessorCount * 16;
    c->reduce_grid = prop.multiProcessorCount * 16;
    *out = c;
    return HH_OK;
}

int hhr_destroy(hhr_ctx *c)
{
    if (!c) return HH_E_ARG;
    (void)hipSetDevice(c->device);
    r_free(c);
    delete c;
    return HH_OK;
}

// Integrate every cluster to tend (a multiple of its dt_max).  Steps are
// launched in batches of batch_steps (captured once into a hipGraph when
// use_graph != 0); after each batch the number of clusters still active is
// read back.  Extra steps after the last cluster finishes are no-ops.
// With force timing in mode 1, steps are launched directly (use_graph
// ignored); mode 2 times graph replays (and direct launches) from device
// clock stamps.
int hhr_run(hhr_ctx *c, double tend, int use_graph, int batch_steps)
{
    if (!c || batch_steps <= 0) return HH_E_ARG;
    if (c->timing == 1) use_graph = 0;
    RCHK(c, hipSetDevice(c->device));
    // set tend for every cluster
    std::vector<RClu> hc(c->nclu);
    RCHK(c, hipMemcpy(hc.data(), c->clu, sizeof(RClu) * c->nclu, hipMemcpyDeviceToHost));
    for (auto &h : hc) h.tend = tend;
    RCHK(c, hipMemcpy(c->clu, hc.data(), sizeof(RClu) * c->nclu, hipMemcpyHostToDevice));
    if (c->timing && c->steplog_cap != batch_steps) {
        r_free_events(c);
        r_drop_graph(c);
        if (c->timing == 1) {
            c->kev.assign(4 * batch_steps, nullptr);
            for (auto &e : c->kev) RCHK(c, hipEventCreate(&e));
        } else {
            RCHK(c, hipMalloc(&c->stamp, sizeof(long long) * (4 * (size_t)batch_steps + 1)));
            RCHK(c, hipMemcpy((char *)c->glob + offsetof(RGlob, stamp), &c->stamp, sizeof(c->stamp),
                              hipMemcpyHostToDevice));
        }
        RCHK(c, hipMalloc(&c->steplog, sizeof(int) * batch_steps));
        c->steplog_cap = batch_steps;
    }
    if (use_graph && (!c->gexec || c->graph_steps != batch_steps)) {
        if (c->gexec) RCHK(c, hipGraphExecDestroy(c->gexec));
        if (c->graph) RCHK(c, hipGraphDestroy(c->graph));
        c->gexec = nullptr;
        c->graph = nullptr;
        RCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int s = 0; s < batch_steps && e == hipSuccess; s++) e = r_step(c, s);
        if (e == hipSuccess) e = r_stamp_end(c);
        hipGraph_t gph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(c->stream, &gph);
        if (e != hipSuccess) return rfail(c, HH_E_HIP, std::string("step capture: ") + hipGetErrorString(e));
        if (e2 != hipSuccess) return rfail(c, HH_E_HIP, std::string("end capture: ") + hipGetErrorString(e2));
        c->graph = gph;
        RCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->graph_steps = batch_steps;
    }
    RGlob g;
    long long launched = 0;
    RCHK(c, hipEventRecord(c->ev0, c->stream));
    for (;;) {
        if (use_graph) {
            RCHK(c, hipGraphLaunch(c->gexec, c->stream));
        } else {
            for (int s = 0; s < batch_steps; s++) {
                const hipError_t e = r_step(c, s);
                if (e != hipSuccess) return rfail(c, HH_E_HIP, std::string("step launch: ") + hipGetErrorString(e));
            }
            const hipError_t e = r_stamp_end(c);
            if (e != hipSuccess) return rfail(c, HH_E_HIP, std::string("stamp launch: ") + hipGetErrorString(e));
        }
        c->steps_launched += batch_steps;
        launched += batch_steps;
        RCHK(c, hipMemcpyAsync(&g, c->glob, sizeof(RGlob), hipMemcpyDeviceToHost, c->stream));
        RCHK(c, hipStreamSynchronize(c->stream));
        if (c->timing) {
            std::vector<int> na(batch_steps);
            RCHK(c, hipMemcpy(na.data(), c->steplog, sizeof(int) * batch_steps, hipMemcpyDeviceToHost));
            std::vector<long long> ts;
            if (c->timing == 2) {
                ts.resize(4 * (size_t)batch_steps + 1);
                RCHK(c, hipMemcpy(ts.data(), c->stamp, sizeof(long long) * ts.size(), hipMemcpyDeviceToHost));
            }
            for (int s = 0; s < batch_steps; s++) {
                float t01 = 0.f, t12 = 0.f, t23 = 0.f;
                if (c->timing == 1) {
                    RCHK(c, hipEventElapsedTime(&t01, c->kev[4 * s], c->kev[4 * s + 1]));
                    RCHK(c, hipEventElapsedTime(&t12, c->kev[4 * s + 1], c->kev[4 * s + 2]));
                    RCHK(c, hipEventElapsedTime(&t23, c->kev[4 * s + 2], c->kev[4 * s + 3]));
                } else {
                    // step s of the batch is step g.step - batch_steps + s, whose
                    // stamps sit at that step modulo the batch size; the next
                    // step's first stamp (or the batch's end) closes it
                    const long long st = (long long)g.step - batch_steps + s;
                    const size_t k = (size_t)(st % batch_steps);
                    const long long t_end = s + 1 < batch_steps ? ts[(size_t)((st + 1) % batch_steps) * 4]
                                                                : ts[4 * (size_t)batch_steps];
                    const double ms_per_tick = 1.0 / c->wall_khz;
                    t01 = (float)((ts[k * 4 + 2] - ts[k * 4]) * ms_per_tick);
                    t12 = (float)((ts[k * 4 + 3] - ts[k * 4 + 2]) * ms_per_tick);
                    t23 = (float)((t_end - ts[k * 4 + 3]) * ms_per_tick);
                }
                c->force_ms += t12;
                // steps batch_steps before g.step wrote their counts at (step % cap)
                const long long st = (long long)g.step - batch_steps + s;
                const int nact = na[(size_t)(st % batch_steps)];
                if (nact > 0) {
                    c->log_nact.push_back(nact);
                    c->log_force.push_back(t12);
                    c->log_other.push_back(t01 + t23);
                }
            }
            c->force_launches += batch_steps;
        }
        if (g.error == 1) return rfail(c, HH_E_STATE, "a cluster's block time stopped advancing (step-size collapse)");
        if (g.error == 2) return rfail(c, HH_E_STATE, "a particle state became non-finite");
        if (g.nactive == 0) break;
        if (c->step_limit > 0 && launched >= c->step_limit)
            return rfail(c, HH_E_LIMIT, "stopped at the step limit (" + std::to_string(launched) +
                                            " block steps launched in this call)");
    }
    RCHK(c, hipEventRecord(c->ev1, c->stream));
    RCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    RCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_run_ms = ms;
    return HH_OK;
}

int hhr_stats(hhr_ctx *c, int k, long long *num_step, long long *num_bstep, double *tsys, long long *dtmin_hits)
{
    if (!c || k < 0 || k >= c->nclu) return HH_E_ARG;
    RCHK(c, hipSetDevice(c->device));
    RClu h;
    RCHK(c, hipMemcpy(&h, c->clu + k, sizeof(RClu), hipMemcpyDeviceToHost));
    if (num_step) *num_step = h.num_step;
    if (num_bstep) *num_bstep = h.num_bstep;
    if (tsys) *tsys = h.tsys;
    if (dtmin_hits) *dtmin_hits = h.dtmin_hits;
    return HH_OK;
}

double hhr_last_run_ms(const hhr_ctx *c) { return c ? c->last_run_ms : -1.0; }

int hhr_set_plan(hhr_ctx *c, int jpar_max, int smax, int task_target, int jpar_target)
{
    if (!c) return HH_E_ARG;
    // j-parallel groups are for short lists: at most R_JPAR_MAX / 2 groups
    // per cluster keeps every step's task count far inside the task fields
    if (jpar_max < 0 || jpar_max > R_JPAR_MAX || smax < 1 || task_target < 1 || task_target > R_TASK_MAX ||
        jpar_target < 1 || jpar_target > R_TASK_MAX)
        return rfail(c, HH_E_ARG, "plan parameters out of range");
    RCHK(c, hipSetDevice(c->device));
    RCHK(c, hipStreamSynchronize(c->stream));
    RPlan p;
    p.jpar_max = jpar_max;
    p.smax = smax;
    p.task_target = task_target;
    p.jpar_target = jpar_target;
    // the split-wave thresholds that follow are hhr_set_split_waves' to set
    RCHK(c, hipMemcpy((char *)c->glob + offsetof(RGlob, plan), &p, offsetof(RPlan, wj4_max), hipMemcpyHostToDevice));
    return HH_OK;
}

int hhr_set_split_waves(hhr_ctx *c, int wj4_max, int wj2_max)
{
    if (!c) return HH_E_ARG;
    RCHK(c, hipSetDevice(c->device));
    RCHK(c, hipStreamSynchronize(c->stream));
    const int v[2] = {wj4_max, wj2_max};
    RCHK(c, hipMemcpy((char *)c->glob + offsetof(RGlob, plan) + offsetof(RPlan, wj4_max), v, sizeof(v),
                      hipMemcpyHostToDevice));
    return HH_OK;
}

int hhr_set_fused(hhr_ctx *c, int max_n, int grid, int jpl, int islice_max)
{
    if (!c || max_n < 0 || max_n > R_FUSED_MAX || grid < 0 || grid > 4096 || !(jpl == 0 || jpl == 1 || jpl == 2 || jpl == 4) ||
        islice_max < -1 || islice_max > R_FUSED_MAX)
        return rfail(c, HH_E_ARG, "fused-step parameters out of range");
    RCHK(c, hipSetDevice(c->device));
    RCHK(c, hipStreamSynchronize(c->stream));
    c->fused_max = max_n;
    c->fused_grid = grid ? grid : 512;
    c->fused_jpl = jpl;
    c->fused_islice_max = islice_max;
    // the captured batch holds the launches of the old choice
    if (c->gexec) RCHK(c, hipGraphExecDestroy(c->gexec));
    if (c->graph) RCHK(c, hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_steps = 0;
    return HH_OK;
}

int hhr_set_step_limit(hhr_ctx *c, long long max_block_steps)
{
    if (!c || max_block_steps < 0) return HH_E_ARG;
    c->step_limit = max_block_steps;
    return HH_OK;
}

int hhr_set_timing(hhr_ctx *c, int on)
{
    if (!c || on < 0 || on > 2) return HH_E_ARG;
    if (c->timing != on) {
        RCHK(c, hipSetDevice(c->device));
        RCHK(c, hipStreamSynchronize(c->stream));
        r_free_events(c);
        r_drop_graph(c);
    }
    c->timing = on;
    c->force_ms = 0.0;
    c->force_launches = 0;
    c->log_nact.clear();
    c->log_force.clear();
    c->log_other.clear();
    return HH_OK;
}

long long hhr_step_log(const hhr_ctx *c, long long cap, int *nact, float *force_ms, float *other_ms)
{
    if (!c) return HH_E_ARG;
    const long long n = (long long)c->log_nact.size();
    if (cap > 0 && nact && force_ms && other_ms) {
        const long long m = cap < n ? cap : n;
        for (long long i = 0; i < m; i++) {
            nact[i] = c->log_nact[i];
            force_ms[i] = c->log_force[i];
            other_ms[i] = c->log_other[i];
        }
    }
    return n;
}

int hhr_force_timing(const hhr_ctx *c, double *ms_total, long long *launches)
{
    if (!c || !ms_total || !launches) return HH_E_ARG;
    *ms_total = c->force_ms;
    *launches = c->force_launches;
    return HH_OK;
}
long long hhr_steps_launched(const hhr_ctx *c) { return c ? c->steps_launched : -1; }

// Reference Particle records of cluster k.
int hhr_get_particles(hhr_ctx *c, int k, double *recs)
{
    if (!c || !recs || k < 0 || k >= c->nclu) return HH_E_ARG;
    RCHK(c, hipSetDevice(c->device));
    const int o = c->off[k], nk = c->n[k], rw = c->rw, pw = p_words(c->order);
    std::vector<double> w((size_t)rw * nk);
    for (int q = 0; q < rw; q++)
        RCHK(c, hipMemcpy(w.data() + (size_t)q * nk, c->rp + (size_t)q * c->npad + o, sizeof(double) * nk,
                          hipMemcpyDeviceToHost));
    for (int i = 0; i < nk; i++) {
        double *p = recs + (size_t)i * pw;
        std::memset(p, 0, sizeof(double) * pw);
        auto g = [&](int q) { return w[(size_t)q * nk + i]; };
        p[0] = g(R_ID);
        p[1] = g(3);
        p[2] = g(7);
        p[3] = g(R_DT);
        for (int d = 0; d < 3; d++) {
            p[4 + d] = g(0 + d);
            p[7 + d] = g(4 + d);
            p[10 + d] = g(8 + d);
            p[13 + d] = g(11 + d);
            p[16 + d] = g(14 + d);
            p[19 + d] = g(17 + d);
            if (c->order >= 6) {
                p[22 + d] = g(20 + d);
                p[25 + d] = g(23 + d);
            }
            if (c->order == 8) {
                p[28 + d] = g(R_D6A + d);
                p[31 + d] = g(R_D7A + d);
            }
        }
    }
    return HH_OK;
}

// Total energy of cluster k (calc_energy, SRC/nbodysystem.h:633-650): the
// potential kernel on the resident records, kinetic and potential sums in
// particle order on the host.
int hhr_energy(hhr_ctx *c, int k, double *energy)
{
    if (!c || !energy || k < 0 || k >= c->nclu) return HH_E_ARG;
    RCHK(c, hipSetDevice(c->device));
    const int o = c->off[k], nk = c->n[k];
    // a private tile launch over this cluster's records
    hh_ctx tmp;
    tmp.order = c->order;
    tmp.nbody = nk;
    tmp.variant = c->variant;
    tmp.device = c->device;
    tmp.npad = c->npad;
    tmp.stream = c->stream;
    tmp.d_part = c->part;
    tmp.part_cap = c->part_cap;
    int r;
    const Geo g{1, 1, 0};
    // FP64 potential for the split variant too (the energy diagnostic)
    if (c->variant == HH_EXACT) r = launch_tile<0, EXACT, false>(&tmp, g, c->rp + o, nk, c->eps2[k], c->pot, nullptr);
    else r = launch_tile<0, FP64, false>(&tmp, g, c->rp + o, nk, c->eps2[k], c->pot, nullptr);
    tmp.stream = nullptr;
    tmp.d_part = nullptr;
    if (r != HH_OK) return rfail(c, r, tmp.err);
    std::vector<double> pot(nk), m(nk), v((size_t)3 * nk);
    RCHK(c, hipMemcpyAsync(pot.data(), c->pot, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
    RCHK(c, hipMemcpyAsync(m.data(), c->rp + (size_t)3 * c->npad + o, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
    for (int d = 0; d < 3; d++)
        RCHK(c, hipMemcpyAsync(v.data() + (size_t)d * nk, c->rp + (size_t)(4 + d) * c->npad + o, sizeof(double) * nk,
                               hipMemcpyDeviceToHost, c->stream));
    RCHK(c, hipStreamSynchronize(c->stream));
    double ke = 0.0, pe = 0.0;
    for (int i = 0; i < nk; i++) {
        const double vx = v[i], vy = v[(size_t)nk + i], vz = v[(size_t)2 * nk + i];
        ke += m[i] * (vx * vx + vy * vy + vz * vz);
        pe += m[i] * pot[i];
    }
    ke *= 0.5;
    pe *= 0.5;
    *energy = ke + pe;
    return HH_OK;
}

}  // extern "C"
