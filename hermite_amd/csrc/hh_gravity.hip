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
#include "../../include/hermite_hip.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>

using namespace hh;

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
    unsigned *d_cnt = nullptr;  // per-i-block arrival counters
    int cnt_cap = 0;
    std::atomic<int> dirty_hi{0};
    int tune_kernel = 0, tune_split = 0, tune_ni = 0, tune_fold = 0;
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

#define HIPCHK(c, call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((c), HH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

int hh_particle_words(int order) { return (order == 4 || order == 6 || order == 8) ? p_words(order) : HH_E_ARG; }
int hh_force_words(int order) { return (order == 4 || order == 6 || order == 8) ? f_words(order) : HH_E_ARG; }
int hh_jp_words(int order) { return (order == 4 || order == 6 || order == 8) ? jp_words(order) : HH_E_ARG; }
int hh_pred_words(int order) { return (order == 4 || order == 6 || order == 8) ? pr_words(order) : HH_E_ARG; }

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
    (void)hipFree(c->d_cnt);
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

// Workgroups wanted per force launch: several per CU on 256 CUs.
static constexpr int TARGET_WG = 2048;
// Up to this many j-slices the last workgroup of each i-block adds them;
// beyond it a separate slice_reduce_kernel does, with the whole GPU.
static constexpr int LAST_BLOCK_MAX = 16;

// i-parallel launch of tile_kernel over src slots [0, ni) as i-particles.
template <int ORDER, int MODE, bool POT, int NI, int FOLD>
static int launch_tile_ni(hh_ctx *c, const double *src, int ni, double eps2, double *fout, double *pout)
{
    constexpr int NC = n_acc(ORDER, POT);
    const int bi = NT * NI;
    const int iblocks = (ni + bi - 1) / bi;
    const int ntiles = (c->nbody + TJ - 1) / TJ;
    int nslice;
    if (MODE == EXACT) nslice = 1;
    else if (c->tune_split > 0) nslice = std::min(c->tune_split, ntiles);
    else nslice = std::max(1, std::min(ntiles, (TARGET_WG + iblocks - 1) / iblocks));
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
    a.cnt = c->d_cnt;
    a.last_block = nslice <= LAST_BLOCK_MAX;
    a.fout = fout;
    a.pout = pout;
    if (nslice > 1) {
        if (iblocks > c->cnt_cap) return fail(c, HH_E_STATE, "counter buffer too small");
        const int r = ensure_part(c, (size_t)nslice * NC * a.ipad);
        if (r) return r;
        a.part = c->d_part;
    }
    hipLaunchKernelGGL((tile_kernel<ORDER, MODE, NI, POT, FOLD>), dim3(iblocks, nslice), dim3(NT), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nslice > 1 && !a.last_block) {
        const int nt = ni * NC;
        hipLaunchKernelGGL(slice_reduce_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream,
                           (const double *)a.part, nslice, NC, ORDER == 0 ? 1 : f_words(ORDER), a.ipad, ni, fout, pout);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("force launch: ") + hipGetErrorString(e));
    c->last_kernel = 1;
    c->last_split = nslice;
    c->last_wgs = iblocks * nslice;
    return HH_OK;
}

// Default FP32 partial-sum length of the SPLIT body.
static constexpr int DEFAULT_FOLD = 8;

template <int ORDER, int MODE, bool POT>
static int launch_tile(hh_ctx *c, const double *src, int ni, double eps2, double *fout, double *pout)
{
    if constexpr (MODE == EXACT) {
        return launch_tile_ni<ORDER, MODE, POT, 1, 1>(c, src, ni, eps2, fout, pout);
    } else if constexpr (MODE == FP64) {
        if (c->tune_ni == 2) return launch_tile_ni<ORDER, MODE, POT, 2, 1>(c, src, ni, eps2, fout, pout);
        return launch_tile_ni<ORDER, MODE, POT, 1, 1>(c, src, ni, eps2, fout, pout);
    } else {
        const int fold = c->tune_fold ? c->tune_fold : DEFAULT_FOLD;
        if (c->tune_ni == 2) {
            switch (fold) {
            case 4: return launch_tile_ni<ORDER, MODE, POT, 2, 4>(c, src, ni, eps2, fout, pout);
            case 16: return launch_tile_ni<ORDER, MODE, POT, 2, 16>(c, src, ni, eps2, fout, pout);
            case 32: return launch_tile_ni<ORDER, MODE, POT, 2, 32>(c, src, ni, eps2, fout, pout);
            default: return launch_tile_ni<ORDER, MODE, POT, 2, 8>(c, src, ni, eps2, fout, pout);
            }
        }
        switch (fold) {
        case 4: return launch_tile_ni<ORDER, MODE, POT, 1, 4>(c, src, ni, eps2, fout, pout);
        case 16: return launch_tile_ni<ORDER, MODE, POT, 1, 16>(c, src, ni, eps2, fout, pout);
        case 32: return launch_tile_ni<ORDER, MODE, POT, 1, 32>(c, src, ni, eps2, fout, pout);
        default: return launch_tile_ni<ORDER, MODE, POT, 1, 8>(c, src, ni, eps2, fout, pout);
        }
    }
}

// Largest active count routed to the j-parallel kernel by default.
static constexpr int JPAR_MAX = 64;
static constexpr int TARGET_WG_JPAR = 1024;

// j-parallel launch (jpar_kernel) over [0, nact) as i-particles.
template <int ORDER, int MODE, int GI, bool POT>
static int launch_jpar(hh_ctx *c, int nact, double eps2)
{
    constexpr int V = GI * n_acc(ORDER, POT);
    const int ngroups = (nact + GI - 1) / GI;
    const int ntj = c->npad / NT;
    int nslice = c->tune_split > 0 ? c->tune_split : (TARGET_WG_JPAR + ngroups - 1) / ngroups;
    nslice = std::max(1, std::min(nslice, ntj));
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
    a.cnt = c->d_cnt;
    a.fout = c->d_force;
    a.pout = c->d_pot;
    if (nslice > 1) {
        if (ngroups > c->cnt_cap) return fail(c, HH_E_STATE, "counter buffer too small");
        const int r = ensure_part(c, (size_t)nslice * ngroups * V);
        if (r) return r;
        a.part = c->d_part;
    }
    hipLaunchKernelGGL((jpar_kernel<ORDER, MODE, GI, POT>), dim3(ngroups, nslice), dim3(NT), 0, c->stream, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, HH_E_HIP, std::string("force launch: ") + hipGetErrorString(e));
    c->last_kernel = 2;
    c->last_split = nslice;
    c->last_wgs = ngroups * nslice;
    return HH_OK;
}

template <int ORDER, int MODE, bool POT>
static int launch_jpar_gi(hh_ctx *c, int nact, double eps2)
{
    if (nact == 1) return launch_jpar<ORDER, MODE, 1, POT>(c, nact, eps2);
    if (nact == 2) return launch_jpar<ORDER, MODE, 2, POT>(c, nact, eps2);
    return launch_jpar<ORDER, MODE, 4, POT>(c, nact, eps2);
}

template <int ORDER, int MODE, bool POT>
static int launch_force_mode(hh_ctx *c, int nact, double eps2)
{
    int kernel = c->tune_kernel;
    if (kernel == 0) kernel = nact <= JPAR_MAX ? 2 : 1;
    if constexpr (MODE != EXACT) {
        if (kernel == 2) return launch_jpar_gi<ORDER, MODE, POT>(c, nact, eps2);
    }
    return launch_tile<ORDER, MODE, POT>(c, c->d_jpr, nact, eps2, c->d_force, c->d_pot);
}

template <int ORDER>
static int launch_force_order(hh_ctx *c, int nact, double eps2, bool pot)
{
    switch (c->variant) {
    case HH_FP64:
        return pot ? launch_force_mode<ORDER, FP64, true>(c, nact, eps2) : launch_force_mode<ORDER, FP64, false>(c, nact, eps2);
    case HH_SPLIT:
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
    default: return launch_force_order<8>(c, nact, eps2, pot);
    }
}

static int launch_potential(hh_ctx *c, double eps2)
{
    switch (c->variant) {
    case HH_FP64: return launch_tile<0, FP64, false>(c, c->d_jst, c->nbody, eps2, c->d_pot, nullptr);
    case HH_SPLIT: return launch_tile<0, SPLIT, false>(c, c->d_jst, c->nbody, eps2, c->d_pot, nullptr);
    default: return launch_tile<0, EXACT, false>(c, c->d_jst, c->nbody, eps2, c->d_pot, nullptr);
    }
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
    if (order != 4 && order != 6 && order != 8) return fail(nullptr, HH_E_ARG, "order must be 4, 6 or 8");
    if (nbody <= 0) return fail(nullptr, HH_E_ARG, "nbody must be positive");
    if (variant < HH_FP64 || variant > HH_EXACT) return fail(nullptr, HH_E_ARG, "unknown variant");
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
    c->cnt_cap = (int)np;  // one per i-group of the finest grouping
    if ((e = hipMalloc(&c->d_cnt, sizeof(unsigned) * c->cnt_cap)) != hipSuccess) return bail("counters", e);
    if ((e = hipMemsetAsync(c->d_cnt, 0, sizeof(unsigned) * c->cnt_cap, c->stream)) != hipSuccess) return bail("memset", e);
    if ((e = hipMemsetAsync(c->d_jst, 0, sizeof(double) * W * np, c->stream)) != hipSuccess) return bail("memset", e);
    std::memset(c->h_jp, 0, sizeof(double) * W * nbody);
    switch (order) {
    case 4: e = launch_pad<4>(c); break;
    case 6: e = launch_pad<6>(c); break;
    default: e = launch_pad<8>(c); break;
    }
    if (e != hipSuccess) return bail("pad", e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("sync", e);
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
    int cur = c->dirty_hi.load(std::memory_order_relaxed);
    while (cur < hi && !c->dirty_hi.compare_exchange_weak(cur, hi, std::memory_order_relaxed)) {
    }
    return HH_OK;
}

// set_jp: SRC/hermite6-titan.h:70-97 (field copy into the pinned j-store).
int hh_set_jp(hh_ctx *c, int addr, const double *p)
{
    if (!c || !p || addr < 0 || addr >= c->nbody) return HH_E_ARG;
    const int W = jp_words(c->order);
    double *d = c->h_jp + (size_t)addr * W;
    d[0] = p[4];  d[1] = p[5];  d[2] = p[6];   // pos
    d[3] = p[1];                               // mass
    d[4] = p[7];  d[5] = p[8];  d[6] = p[9];   // vel
    d[7] = p[2];                               // tlast
    for (int w = 0; w < 6; w++) d[8 + w] = p[10 + w];                     // acc, jrk
    if (c->order >= 6) for (int w = 0; w < 6; w++) d[14 + w] = p[16 + w];  // snp, crk
    if (c->order == 8) for (int w = 0; w < 6; w++) d[20 + w] = p[22 + w];  // d4a, d5a
    return hh_mark_dirty(c, 0, addr + 1);
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

int hh_last_geometry(const hh_ctx *c, int *kernel, int *jsplit, int *workgroups)
{
    if (!c) return HH_E_ARG;
    if (kernel) *kernel = c->last_kernel;
    if (jsplit) *jsplit = c->last_split;
    if (workgroups) *workgroups = c->last_wgs;
    return HH_OK;
}

}  // extern "C"
