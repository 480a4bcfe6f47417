// Host side of the exact Debye engine: the C ABI declared in
// include/pripps_debye.h.
//
// A build takes the caller's arrays, validates them, sorts atoms by
// (kind, hydrated) class, centres and converts the coordinates, lays out one
// work queue of (structure, i-tile, j-span, q-strip) items, uploads
// everything with a single copy from pinned staging into a grow-only device
// arena, and runs the pair kernel, the finalize kernel and nothing else.
// Models keep only their q grid and six profiles on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <tuple>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "internal.h"
#include "pripps_debye.h"

namespace pripps {

// fit (fit_kernels.hip)
size_t device_fit_doubles(int nq, int64_t n, int method, int c1_points, int c2_points);
int device_fit(hipStream_t st, double *buf, const double *d_prof, const double *d_q, int nq,
               const double *q, const double *inten, const double *sigma, int64_t n, int method,
               int c1_points, int c2_points, const double *c1b, const double *c2b,
               pripps_fit_result *out, std::string *err);
size_t device_spline_total_doubles(int nq, int64_t n);
int device_spline_total(hipStream_t st, double *buf, const double *d_prof, const double *d_q,
                        int nq, double vs, double cd, const double *qe, int64_t n, double *out,
                        std::string *err);

}  // namespace pripps

using namespace pripps;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

static int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? PRIPPS_ERR_OUT_OF_MEMORY : PRIPPS_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(expr)                                    \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

extern "C" const char *pripps_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *pripps_version(void) { return "pripps_amd 0.3.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// device block pool for model buffers
// ---------------------------------------------------------------------------
// A model holds [6 nq profiles | nq q | 4 nq combine output] doubles for its
// whole life.  Those blocks come from a per-device pool of power-of-two size
// classes (4 KiB and up) carved from 4 MiB slabs; a freed block goes back to
// its class's free list.  Slabs stay with the process, so releasing a model
// never calls hipFree, which would wait for every stream of the device -- in
// the serving loop, for the other lanes' pair kernels.  What the pool holds
// is the high-water mark of live models per class.
namespace {
constexpr int kPoolMinLog = 12;
constexpr int kPoolMaxLog = 46;
constexpr size_t kPoolSlab = (size_t)4 << 20;

struct DevicePool {
    std::mutex mu;
    std::vector<void *> free_list[kPoolMaxLog + 1];
    size_t held = 0;
};

DevicePool &device_pool(int device) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<DevicePool>> pools;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)pools.size() <= device) pools.resize((size_t)device + 1);
    if (!pools[device]) pools[device].reset(new DevicePool());
    return *pools[device];
}

int pool_class(size_t bytes) {
    int c = kPoolMinLog;
    while (c < kPoolMaxLog && ((size_t)1 << c) < bytes) ++c;
    return c;
}
}  // namespace

namespace pripps {
// The caller has `device` current.
hipError_t pool_alloc(int device, size_t bytes, void **out) {
    *out = nullptr;
    const int c = pool_class(bytes);
    if (((size_t)1 << c) < bytes) return hipErrorOutOfMemory;
    DevicePool &P = device_pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    auto &fl = P.free_list[c];
    if (fl.empty()) {
        const size_t bs = (size_t)1 << c;
        const size_t slab = std::max(bs, kPoolSlab);
        void *base = nullptr;
        hipError_t err = hipMalloc(&base, slab);
        if (err != hipSuccess) return err;
        P.held += slab;
        for (size_t off = slab; off > 0; off -= bs) fl.push_back((char *)base + off - bs);
    }
    *out = fl.back();
    fl.pop_back();
    return hipSuccess;
}

void pool_free(int device, void *p, size_t bytes) {
    if (!p) return;
    DevicePool &P = device_pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    P.free_list[pool_class(bytes)].push_back(p);
}

size_t pool_held_bytes(int device) {
    DevicePool &P = device_pool(device);
    std::lock_guard<std::mutex> lk(P.mu);
    return P.held;
}
}  // namespace pripps

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
namespace pripps {
// One persistent launch of the pair kernel plus its reduction, replayable on
// the inputs still resident in the engine arena.
struct LaunchRecord {
    bool fp64 = false, uniform = true;
    int Q = 64, P = 0, blocks = 0, n_strips = 0;
    int64_t n_items = 0;
    int *counter = nullptr;
    TiledArgs targs{};
    FinalizeArgs fargs{};     // recurrence kernels
    ChebReduceArgs cargs{};   // Chebyshev kernel
};

// fill counter (recurrence kernels), pair kernel (bracketed by ev_a/ev_b when
// given), reduction
hipError_t run_launch(const LaunchRecord &r, hipStream_t st, hipEvent_t ev_a, hipEvent_t ev_b) {
    // the Chebyshev kernel leaves its queue counters at zero when it ends, and
    // the upload zeroes them, so only the recurrence kernels need the fill
    hipError_t err = r.Q == kChebQ ? hipSuccess : launch_fill_int(r.counter, 1, 0, st);
    if (err == hipSuccess && ev_a) err = hipEventRecord(ev_a, st);
    if (err == hipSuccess && r.n_items > 0)
        err = r.Q != kChebQ ? launch_tiled(r.fp64, r.Q, r.uniform, r.targs, r.blocks, st)
              : r.fp64      ? launch_cheb64(r.P, r.targs, r.blocks, st)
                            : launch_cheb(r.P, r.targs, r.blocks, st);
    if (err == hipSuccess && ev_b) err = hipEventRecord(ev_b, st);
    if (err == hipSuccess && r.Q == kChebQ) err = launch_cheb_reduce(r.cargs, st);
    else if (err == hipSuccess && r.n_strips > 0) err = launch_finalize(r.fargs, r.n_strips, st);
    return err;
}
}  // namespace pripps

struct pripps_engine {
    std::atomic<int> refs{1};  // the creator's handle plus one per live model
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;  // serialises every submission on `stream` and the arenas
    void *arena = nullptr;
    size_t arena_size = 0;
    std::atomic<uint64_t> arena_pub{0};  // arena_size, readable without `mu` (serving admission)
    size_t arena_prev = 0;               // size of the last arena allocated (growth step)
    void *staging = nullptr;  // pinned (hipHostMalloc)
    size_t staging_size = 0;
    // fit scratch, grow-only: a hipFree per fit call would wait for every
    // stream of the device (the serving lanes' pair kernels included)
    double *fit_buf = nullptr;
    size_t fit_doubles = 0;
    // launches of the most recent build, valid while the arena holds them
    uint64_t build_serial = 0;
    bool last_plain = false;
    std::vector<LaunchRecord> last_launches;
    // those launches captured as one HIP graph (replays without per-kernel
    // timing), for the build `replay_serial`
    hipGraph_t replay_graph = nullptr;
    hipGraphExec_t replay_exec = nullptr;
    uint64_t replay_serial = 0;
    // autotuned plans per size class (pripps_engine_autotune)
    std::mutex tune_mu;
    std::vector<pripps_tuning_entry> tune_table;
};

struct pripps_debye_model {
    pripps_engine *engine = nullptr;
    int32_t nq = 0;
    std::vector<double> q;
    double *dev = nullptr;  // [6 nq profiles | nq q | 4 nq combine output], a pool block
    int device = 0;         // the pool `dev` belongs to
    pripps_build_stats stats{};
    uint64_t build_serial = 0;
    double *d_prof() const { return dev; }
    double *d_q() const { return dev + 6 * (size_t)nq; }
    double *d_out() const { return dev + 7 * (size_t)nq; }
};

static void engine_release(pripps_engine *e);

static size_t model_bytes(int32_t nq) { return std::max<size_t>(1, (size_t)11 * nq) * sizeof(double); }

// The model's device block, from the pool of its engine's device (current).
static hipError_t model_alloc(pripps_debye_model *m) {
    m->device = m->engine->device;
    void *p = nullptr;
    hipError_t err = pool_alloc(m->device, model_bytes(m->nq), &p);
    m->dev = (double *)p;
    return err;
}

static void model_release_block(pripps_debye_model *m) {
    if (m->dev) pool_free(m->device, m->dev, model_bytes(m->nq));
    m->dev = nullptr;
}

// Caller holds e->mu with e's device current.
static int ensure_arena(pripps_engine *e, size_t bytes) {
    if (bytes <= e->arena_size) return 0;
    if (e->arena) {
        hipError_t err = hipFree(e->arena);
        e->arena = nullptr;
        e->arena_size = 0;
        e->arena_pub = 0;
        if (err != hipSuccess) return hip_fail(err, "hipFree(arena)");
    }
    // a quarter of headroom over the previous size, in 2 MiB steps, so that a
    // slowly growing stream of builds reallocates (and waits for the device
    // in hipFree) only now and then
    const size_t step = (size_t)2 << 20;
    size_t sz = std::max(bytes, e->arena_prev + e->arena_prev / 4);
    sz = (sz + step - 1) / step * step;
    hipError_t err = hipMalloc(&e->arena, sz);
    if (err != hipSuccess && sz > bytes) {
        sz = bytes;  // the headroom does not fit: take what this build needs
        err = hipMalloc(&e->arena, sz);
    }
    if (err != hipSuccess) return hip_fail(err, "hipMalloc(arena)");
    e->arena_prev = sz;
    e->arena_size = sz;
    e->arena_pub = sz;
    return 0;
}

static int ensure_staging(pripps_engine *e, size_t bytes) {
    if (bytes <= e->staging_size) return 0;
    if (e->staging) {
        hipError_t err = hipHostFree(e->staging);
        e->staging = nullptr;
        e->staging_size = 0;
        if (err != hipSuccess) return hip_fail(err, "hipHostFree(staging)");
    }
    size_t sz = std::max(bytes, (size_t)16 << 20);
    hipError_t err = hipHostMalloc(&e->staging, sz, hipHostMallocDefault);
    if (err != hipSuccess) return hip_fail(err, "hipHostMalloc(staging)");
    e->staging_size = sz;
    return 0;
}

extern "C" int pripps_engine_create(int device, pripps_engine **out) {
    if (!out) return fail(PRIPPS_ERR_USAGE, "pripps_engine_create: out is NULL");
    *out = nullptr;
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(PRIPPS_ERR_USAGE, "pripps_engine_create: device " + std::to_string(device) +
                                          " out of range (" + std::to_string(count) + " devices)");
    HIPCHK(hipSetDevice(device));
    auto *e = new pripps_engine();
    e->device = device;
    hipError_t err = hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&e->ev0);
    if (err == hipSuccess) err = hipEventCreate(&e->ev1);
    if (err != hipSuccess) {
        engine_release(e);
        return hip_fail(err, "pripps_engine_create");
    }
    *out = e;
    return PRIPPS_OK;
}

// Caller holds e->mu with e's device current.
static int ensure_fit(pripps_engine *e, size_t doubles) {
    if (doubles <= e->fit_doubles) return 0;
    if (e->fit_buf) {
        hipError_t err = hipFree(e->fit_buf);
        e->fit_buf = nullptr;
        e->fit_doubles = 0;
        if (err != hipSuccess) return hip_fail(err, "hipFree(fit scratch)");
    }
    const size_t sz = std::max(doubles, (size_t)1 << 16);
    hipError_t err = hipMalloc((void **)&e->fit_buf, sz * sizeof(double));
    if (err != hipSuccess) return hip_fail(err, "hipMalloc(fit scratch)");
    e->fit_doubles = sz;
    return 0;
}

// Fit scratch above this many doubles (64 MiB) is given back after the call
// that needed it: a large grid fit should not keep that memory on the engine
// (and out of a serving budget) for the engine's lifetime.  device_fit and
// device_spline_total synchronise the stream before they return, so nothing
// in flight still uses the buffer.  Caller holds e->mu.
constexpr size_t kFitKeepDoubles = (size_t)8 << 20;

static void trim_fit(pripps_engine *e) {
    if (e->fit_doubles <= kFitKeepDoubles) return;
    (void)hipFree(e->fit_buf);
    e->fit_buf = nullptr;
    e->fit_doubles = 0;
}

static void engine_release(pripps_engine *e) {
    if (e->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->fit_buf) (void)hipFree(e->fit_buf);
    if (e->arena) (void)hipFree(e->arena);
    if (e->staging) (void)hipHostFree(e->staging);
    if (e->replay_exec) (void)hipGraphExecDestroy(e->replay_exec);
    if (e->replay_graph) (void)hipGraphDestroy(e->replay_graph);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int pripps_engine_destroy(pripps_engine *e) {
    if (e) engine_release(e);  // models still alive keep the engine until freed
    return PRIPPS_OK;
}

extern "C" int pripps_engine_memory(pripps_engine *e, uint64_t *free_b, uint64_t *total_b) {
    if (!e) return fail(PRIPPS_ERR_USAGE, "pripps_engine_memory: engine is NULL");
    HIPCHK(hipSetDevice(e->device));
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return PRIPPS_OK;
}

// ---------------------------------------------------------------------------
// host preparation of one structure
// ---------------------------------------------------------------------------
namespace {

struct Prepared {
    int32_t n = 0, nq = 0, K = 0;
    bool has_e = false, has_h = false;
    bool uniform = false;
    bool small_h = false;   // dq * r_max / (4 pi) <= 1/8
    bool sin_direct = false;  // q_max * r_max / (2 pi) < 255 revolutions
    double q0 = 0, dq = 0;
    std::vector<double> q;
    std::vector<double> pos;  // [n][4] centred x y z w, sorted by class
    std::vector<double> ffa, ffe, ffw, kind_n, kind_w;
    double sum_w2 = 0;
    int32_t n_classes = 0;
    int32_t cls_start[kMaxClasses + 1]{};
    int8_t cls_kind[kMaxClasses]{};
    int8_t cls_wet[kMaxClasses]{};
    std::vector<int4> tiles;  // (i0, count, kind, 0)
    double pair_q_terms = 0;
    double rmax = 0;          // upper bound of every pair distance (A)
};

bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int prepare(const pripps_debye_input *in, Prepared &P) {
    if (!in) return fail(PRIPPS_ERR_USAGE, "input descriptor is NULL");
    const int64_t n = in->n_atoms;
    if (n < 0) return fail(PRIPPS_ERR_VALIDATION, "n_atoms must be >= 0");
    if (n > (int64_t(1) << 30)) return fail(PRIPPS_ERR_VALIDATION, "n_atoms exceeds 2^30");
    if (in->nq < 0) return fail(PRIPPS_ERR_VALIDATION, "nq must be >= 0");
    if (in->nq > (1 << 24)) return fail(PRIPPS_ERR_VALIDATION, "nq exceeds 2^24");
    if (n > 0 && (!in->x || !in->y || !in->z || !in->kind))
        return fail(PRIPPS_ERR_USAGE, "positions and kind arrays must be non-NULL");
    if (in->nq > 0 && !in->q) return fail(PRIPPS_ERR_USAGE, "q array is NULL");
    if (in->n_kinds < 1) return fail(PRIPPS_ERR_VALIDATION, "n_kinds must be >= 1");
    if (!in->atom_ff && in->nq > 0) return fail(PRIPPS_ERR_USAGE, "atom_ff is NULL");
    const bool has_w = in->hydration_weight != nullptr, has_wf = in->water_ff != nullptr;
    if (has_w != has_wf)
        return fail(PRIPPS_ERR_USAGE,
                    "hydration needs both water_ff and hydration_weight (or neither)");
    P.n = (int32_t)n;
    P.nq = in->nq;
    P.has_e = in->excluded_ff != nullptr;
    P.has_h = has_w;
    const int nq = P.nq;
    P.q.assign(in->q, in->q + nq);
    for (int m = 0; m < nq; ++m)
        if (!std::isfinite(P.q[m]) || P.q[m] < 0.0)
            return fail(PRIPPS_ERR_VALIDATION, "q values must be finite and >= 0 (q[" +
                                                   std::to_string(m) + "] = " +
                                                   std::to_string(P.q[m]) + ")");
    // used kinds -> dense ids
    std::vector<int32_t> dense(in->n_kinds, -1);
    std::vector<int32_t> used;
    for (int64_t i = 0; i < n; ++i) {
        int32_t k = in->kind[i];
        if (k < 0 || k >= in->n_kinds)
            return fail(PRIPPS_ERR_VALIDATION, "kind id " + std::to_string(k) + " of atom " +
                                                   std::to_string(i) + " outside [0, n_kinds)");
        if (dense[k] < 0) {
            dense[k] = 0;
            used.push_back(k);
        }
    }
    std::sort(used.begin(), used.end());
    if ((int)used.size() > kMaxKinds)
        return fail(PRIPPS_ERR_VALIDATION, "structure uses " + std::to_string(used.size()) +
                                               " atom kinds; the engine supports at most " +
                                               std::to_string(kMaxKinds));
    for (size_t r = 0; r < used.size(); ++r) dense[used[r]] = (int32_t)r;
    const int K = std::max<int>(1, (int)used.size());
    P.K = K;
    P.ffa.assign((size_t)K * nq, 0.0);
    if (P.has_e) P.ffe.assign((size_t)K * nq, 0.0);
    for (size_t r = 0; r < used.size(); ++r) {
        const double *src = in->atom_ff + (size_t)used[r] * nq;
        if (!finite_all(src, nq))
            return fail(PRIPPS_ERR_VALIDATION, "atom form factors must be finite");
        std::copy(src, src + nq, P.ffa.begin() + r * nq);
        if (P.has_e) {
            const double *se = in->excluded_ff + (size_t)used[r] * nq;
            if (!finite_all(se, nq))
                return fail(PRIPPS_ERR_VALIDATION, "excluded form factors must be finite");
            std::copy(se, se + nq, P.ffe.begin() + r * nq);
        }
    }
    if (P.has_h) {
        if (!finite_all(in->water_ff, nq))
            return fail(PRIPPS_ERR_VALIDATION, "water form factor must be finite");
        P.ffw.assign(in->water_ff, in->water_ff + nq);
        if (!finite_all(in->hydration_weight, (size_t)n))
            return fail(PRIPPS_ERR_VALIDATION, "hydration weights must be finite");
    }
    // centroid
    double cx = 0, cy = 0, cz = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(in->x[i]) || !std::isfinite(in->y[i]) || !std::isfinite(in->z[i]))
            return fail(PRIPPS_ERR_VALIDATION, "coordinates must be finite (atom " +
                                                   std::to_string(i) + ")");
        cx += in->x[i];
        cy += in->y[i];
        cz += in->z[i];
    }
    if (n > 0) {
        cx /= (double)n;
        cy /= (double)n;
        cz /= (double)n;
    }
    // classes: every hydrated class (by dense kind) before every dry one.
    // Pairs are visited as i < j, so a hydrated j atom then only meets the
    // hydrated atoms before it: the pair terms that need the w_j sums drop
    // from about f (the hydrated fraction of atoms) to f^2 / 2 of all pairs.
    // The counting sort keeps the input order inside a class.
    const int C = 2 * K;
    std::vector<int64_t> count(C + 1, 0);
    auto cls_of = [&](int64_t i) {
        const double w = P.has_h ? in->hydration_weight[i] : 0.0;
        return (w != 0.0 ? 0 : K) + dense[in->kind[i]];
    };
    for (int64_t i = 0; i < n; ++i) count[cls_of(i) + 1]++;
    for (int c = 0; c < C; ++c) count[c + 1] += count[c];
    P.pos.assign((size_t)n * 4, 0.0);
    P.kind_n.assign(K, 0.0);
    P.kind_w.assign(K, 0.0);
    std::vector<int64_t> fillp(count.begin(), count.end() - 1);
    for (int64_t i = 0; i < n; ++i) {
        const int c = cls_of(i);
        const int64_t d = fillp[c]++;
        const double w = P.has_h ? in->hydration_weight[i] : 0.0;
        P.pos[4 * d + 0] = in->x[i] - cx;
        P.pos[4 * d + 1] = in->y[i] - cy;
        P.pos[4 * d + 2] = in->z[i] - cz;
        P.pos[4 * d + 3] = w;
        P.kind_n[c % K] += 1.0;
        P.kind_w[c % K] += w;
        P.sum_w2 += w * w;
    }
    P.n_classes = C;
    for (int c = 0; c <= C; ++c) P.cls_start[c] = (int32_t)count[c];
    for (int c = 0; c < C; ++c) {
        P.cls_kind[c] = (int8_t)(c % K);
        P.cls_wet[c] = (int8_t)(c < K ? 1 : 0);
    }
    // class-pure i-tiles of up to 64 atoms
    for (int c = 0; c < C; ++c) {
        const int32_t b = P.cls_start[c], e = P.cls_start[c + 1];
        for (int32_t i0 = b; i0 < e; i0 += 64)
            P.tiles.push_back(make_int4(i0, std::min(64, e - i0), c % K, 0));
    }
    // uniform grid detection (q_m = q0 + m dq to a few ulps)
    if (nq >= 1) {
        P.q0 = P.q[0];
        P.dq = nq >= 2 ? (P.q[nq - 1] - P.q[0]) / (double)(nq - 1) : 0.0;
        double qmax = 0;
        for (double v : P.q) qmax = std::max(qmax, std::fabs(v));
        double dev = 0;
        for (int m = 0; m < nq; ++m) dev = std::max(dev, std::fabs(P.q[m] - (P.q0 + m * P.dq)));
        P.uniform = dev <= 64.0 * 2.220446049250313e-16 * std::max(qmax, 1e-300);
    }
    // largest pair distance is at most twice the largest distance from the centroid
    double rmax2 = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double *p4 = &P.pos[4 * i];
        rmax2 = std::max(rmax2, p4[0] * p4[0] + p4[1] * p4[1] + p4[2] * p4[2]);
    }
    const double rmax = 2.0 * std::sqrt(rmax2) * (1.0 + 1e-6) + 1e-6;
    P.rmax = rmax;
    P.small_h = std::fabs(P.dq) * rmax / (4.0 * M_PI) <= 0.125;
    double qabs = 0;
    for (double v : P.q) qabs = std::max(qabs, std::fabs(v));
    // v_sin_f32 takes |x| <= 256 revolutions; 255 leaves room for the FP32
    // rounding of q / 2 pi and r
    P.sin_direct = qabs * rmax / (2.0 * M_PI) < 255.0;
    P.pair_q_terms = (double)n * (double)(n - 1) / 2.0 * (double)nq;
    return PRIPPS_OK;
}

struct Layout {
    size_t off = 0;
    size_t take(size_t bytes) {
        off = (off + 255) & ~size_t(255);
        size_t at = off;
        off += bytes;
        return at;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// per-structure plan: q points per lane, j-span, pair block, inner-loop variant
// ---------------------------------------------------------------------------
namespace {

struct StructPlan {
    int Q = 64, span = 256, block = 64, variant = 3;
    bool span_set = false;  // span from the caller's tuning or the tuning table
    // Chebyshev plans (Q == kChebQ): expansion order, strips, and the groups
    // of one or two consecutive strips that share a half-width (one work
    // item covers a group)
    int P = 0;
    std::vector<int32_t> strip_first;  // [n_strips + 1]
    std::vector<double> strip_geo;     // [n_strips][2]: centre, half-width
    std::vector<int32_t> group_first;  // [n_groups + 1]: first strip of each group
    int n_strips(int nq) const {
        return Q == kChebQ ? (int)strip_first.size() - 1 : (nq + Q - 1) / Q;
    }
    // work units per (tile, j-span): strip groups, or strips
    int n_units(int nq) const {
        return Q == kChebQ ? (int)group_first.size() - 1 : n_strips(nq);
    }
    int unit_first(int u) const { return Q == kChebQ ? group_first[u] : u; }
    int unit_strips(int u) const { return Q == kChebQ ? group_first[u + 1] - group_first[u] : 1; }
};

// Q: q-points per lane strip of the recurrence kernels (32, 64, 128), or
// kChebQ for the Chebyshev kernels (FP32 or FP64, any grid).
int valid_q(int precision, bool uniform, int Q) {
    if (precision == PRIPPS_FP64) return (Q == 32 || (Q == 64 && uniform) || Q == kChebQ) ? Q : 0;
    return (Q == 32 || Q == 64 || (Q == 128 && uniform) || Q == kChebQ) ? Q : 0;
}

// Strips of a Chebyshev plan: runs of up to kChebQ consecutive q-points
// whose half-width H keeps beta = H r_max within what order P admits.  The
// expansion's error is about uniform in absolute terms across a strip, while
// the profiles divide the pair sum by q, so a strip also keeps
// H <= kChebLowQ * (its smallest non-zero q): that bounds the amplification
// of the error at the strip's low end (q = 0 itself takes the closed form).
// Consecutive strips then pair up into groups that share the larger H (the
// pair's Bessel weights depend on beta alone, so one recurrence serves both)
// where that H still meets both strips' low-q bound.  Each built order is
// tried with greedy strips and the one with the fewest instructions per pair
// over the grid wins: about 30 + 2 P per group (geometry, the weights'
// recurrence, its normalisation) and 29 + P per strip (its sine and cosine,
// its P sums).  FP64 plans keep one strip per group (two strips' FP64 sums
// and weights would not fit the kernel's registers at two waves per SIMD):
// about 41 + 2 P per group and 42 + P per strip, FP64 instructions, with the
// orders whose truncation stays below 1e-15.
constexpr double kChebLowQ = 128.0;

void cheb_strips(const Prepared &p, StructPlan &sp, bool fp64) {
    static double limit[kChebNumOrders] = {0}, limit64[kCheb64NumOrders] = {0};
    static std::once_flag once;
    std::call_once(once, [] {
        for (int k = 0; k < kChebNumOrders; ++k) limit[k] = cheb_beta_limit(kChebOrders[k]);
        for (int k = 0; k < kCheb64NumOrders; ++k) limit64[k] = cheb64_beta_limit(kCheb64Orders[k]);
    });
    const int *orders = fp64 ? kCheb64Orders : kChebOrders;
    const double *limits = fp64 ? limit64 : limit;
    const int n_orders = fp64 ? kCheb64NumOrders : kChebNumOrders;
    const int nq = p.nq;
    // the distance bound rounded up to a step of 2^(1/8): structures of about
    // the same extent (jittered frames of one topology, say) get one plan
    const double rplan = std::exp2(std::ceil(8.0 * std::log2(std::max(p.rmax, 1e-6))) / 8.0);
    double best_cost = 0;
    for (int k = 0; k < n_orders; ++k) {
        const double hmax = limits[k] / rplan;
        std::vector<int32_t> first{0};
        std::vector<double> geo, nzs;
        int m0 = 0;
        while (m0 < nq) {
            double lo = p.q[m0], hi = p.q[m0];
            double nz = p.q[m0] > 0.0 ? p.q[m0] : HUGE_VAL;  // smallest non-zero q
            int m1 = m0 + 1;
            while (m1 < nq && m1 - m0 < kChebQ) {
                const double l2 = std::min(lo, p.q[m1]), h2 = std::max(hi, p.q[m1]);
                const double nz2 = p.q[m1] > 0.0 ? std::min(nz, p.q[m1]) : nz;
                const double half = 0.5 * (h2 - l2);
                if (half > hmax || (nz2 < HUGE_VAL && half > kChebLowQ * nz2)) break;
                lo = l2;
                hi = h2;
                nz = nz2;
                ++m1;
            }
            first.push_back(m1);
            geo.push_back(0.5 * (lo + hi));
            geo.push_back(0.5 * (hi - lo));
            nzs.push_back(nz);
            m0 = m1;
        }
        const int ns = (int)first.size() - 1;
        std::vector<int32_t> groups{0};
        for (int st = 0; st < ns;) {
            if (st + 1 < ns && !fp64) {
                const double hg = std::max(geo[2 * st + 1], geo[2 * st + 3]);
                const double nzg = std::min(nzs[st], nzs[st + 1]);
                if (nzg == HUGE_VAL || hg <= kChebLowQ * nzg) {
                    geo[2 * st + 1] = geo[2 * st + 3] = hg;
                    st += 2;
                    groups.push_back(st);
                    continue;
                }
            }
            st += 1;
            groups.push_back(st);
        }
        const double cost =
            fp64 ? (double)(groups.size() - 1) * (41.0 + 2.0 * orders[k]) + (double)ns * (42.0 + orders[k])
                 : (double)(groups.size() - 1) * (30.0 + 2.0 * orders[k]) + (double)ns * (29.0 + orders[k]);
        if (k == 0 || cost < best_cost) {
            best_cost = cost;
            sp.P = orders[k];
            sp.strip_first = std::move(first);
            sp.strip_geo = std::move(geo);
            sp.group_first = std::move(groups);
        }
    }
}

// From 3e7 pair-q terms (N^2 nq / 2) on, both precisions take the Chebyshev
// kernels on any grid (kChebQ).  Measured against the best recurrence plan
// (pair kernel, one box), FP32: 1,000 atoms x 500 q 0.066 against 0.101 ms,
// 2,000 x 200 0.088 against 0.127, 10,000 x 500 0.634 against 1.706, 30,000
// x 800 6.8 against 22.9, and on a log grid 10,000 x 1,000 1.30 against 7.79
// ms; FP64: 1,000 x 500 0.055 against 0.133 ms, 10,000 x 500 0.96 against
// 4.32 ms.  Only the smallest structures (300 x 200: FP32 0.077 against
// 0.054 ms, FP64 0.061 against 0.051) are faster with the short recurrence
// strips.
//
// Below that, a small structure alone is latency-bound: uniform FP32 grids
// of more than 32 q-points take 64-point strips, everything else 32 (a
// direct sine per term on non-uniform grids).  The longer recurrence strips
// (FP32 128, FP64 64) only won past the threshold; they stay for explicit
// plans.
constexpr double kChebTerms = 3e7;
int default_q(int precision, bool uniform, int nq, int n_atoms) {
    const double terms = 0.5 * (double)n_atoms * (double)n_atoms * (double)nq;
    if (terms >= kChebTerms) return kChebQ;
    if (!uniform || nq <= 32 || precision == PRIPPS_FP64) return 32;
    return 64;
}

// j atoms per work item: about 16 items per wave slot of the GPU when the
// structure runs alone, 64..8192, a multiple of 64 (256 and up from mid-size
// structures on).  It depends on the structure only, so a structure is cut
// into the same items alone or in any batch, and its profiles are bitwise
// the same either way.
//
// Chebyshev plans (Q = kChebQ) cut their work into units themselves (span
// 0, plan_cheb_group); the span below, about three units per wave slot of
// the GPU (64..8192), is only the base that a tuning-table shift scales.
int default_span(const Prepared &p, int Q, int n_cu, int n_units = 0) {
    const double strips = n_units > 0 ? (double)n_units : (double)((p.nq + Q - 1) / Q);
    const double work = (double)p.tiles.size() * 0.5 * p.n * strips;
    if (Q == kChebQ) {
        const double items = 3.0 * 12.0 * std::max(1, n_cu);
        const int span = (int)std::ceil(work / items);
        return std::max(64, std::min(8192, (span + 63) / 64 * 64));
    }
    const double slots = std::max(1, n_cu) * 8.0;
    const int span = (int)(work / (16.0 * slots));
    const int floor = p.n >= 4096 ? 256 : 64;
    return std::max(floor, std::min(8192, (span + 63) / 64 * 64));
}

// j atoms summed in FP32 per lane before the FP64 carry.  Longer blocks are
// ~3 % faster, but lattice-placed points (grid solvent dummies) repeat exact
// pair distances, and the FP32 rounding of such repeated terms adds up
// coherently within a block: on a grid hydration shell at its first minimum
// the error was 1.1e-5 at 256, 4.5e-6 at 64 and 1.8e-6 at 8.
constexpr int kDefaultBlock = 64;
// Chebyshev plans sum each coefficient over the chunk in FP32 and evaluate
// the sums at every q-point of the strip, so their rounding counts about
// sqrt(P) times more than a sum per q-point; the carry per chunk (lane swaps
// and FP64 adds) is cheap.  On 512 mixed globules the worst error against
// FP64 was 2.0e-6 at 32, 2.4e-6 at 48 and 3.6e-6 at 64 (batch pass 118,
// 114 and 112 ms).
constexpr int kChebBlock = 48;

int shifted_span(int base, int shift) {
    double s = std::ldexp((double)base, shift);
    s = std::max(64.0, std::min(16384.0, s));
    return ((int)s + 63) / 64 * 64;
}

int valid_variant(int precision, bool uniform, int Q, int v) {
    if (precision != PRIPPS_FP32_COMPENSATED || !uniform) return 3;
    if ((v == 1 || v == 2) && Q == 64) return v;
    if (v == 4 && (Q == 32 || Q == 64)) return 4;
    return 3;
}

int log2_atoms_of(int64_t n) { return n <= 1 ? 0 : (int)std::lround(std::log2((double)n)); }

int log2_nq_of(int nq) {
    int b = 0;
    while ((1 << b) < nq && b < 30) ++b;
    return b;
}

// Structures of one kernel shape (Q, uniform grid, variant) share one
// persistent launch.
struct LaunchGroup {
    bool uniform = true;
    bool fp64 = false;  // Chebyshev plans: the FP64 kernel (one strip per group)
    int Q = 64, variant = 3, P = 0;
    std::vector<int> members;
    std::vector<Item> items;
    int64_t n_slots = 0;  // partial slots (recurrence kernels): one per item
    // Chebyshev plans: the queue of units, and the reduction of their records
    std::vector<Unit> units;
    int64_t unit_len = 0;  // longest unit of a member, in j atoms of a two-strip group
    int RS = 0;
    int64_t n_records = 0;
    std::vector<int32_t> chunk_recs, chunk_first;
    std::vector<int4> sb_a, sb_b, st;
    size_t o_units = 0, o_chunk_recs = 0, o_chunk_first = 0, o_sb_a = 0, o_sb_b = 0, o_st = 0;
    size_t o_records = 0, o_chunk_tot = 0, o_sb_out = 0;
    std::vector<int32_t> csr_ptr, csr_items, strip_struct, strip_index, stage_ptr;
    size_t o_items = 0, o_csr_ptr = 0, o_csr_items = 0, o_ss = 0, o_si = 0, o_sp = 0;
    size_t o_counter = 0, o_partials = 0, o_stage = 0;
};

// Everything a submission needs: per-structure plans, launch groups and the
// arena layout.  [0, upload_bytes) of the arena mirrors the pinned staging
// buffer and arrives with one copy; the rest is device-only scratch.
struct BuildPlan {
    bool fp64 = false, plain = false;
    int ns = 0;
    std::vector<StructPlan> sp;
    std::vector<LaunchGroup> groups;
    std::vector<size_t> o_atoms, o_ffa, o_ffe, o_ffw, o_kn, o_kw, o_tiles, o_pp, o_sf, o_sg;
    size_t o_structs = 0, o_user = 0, upload_bytes = 0, total_bytes = 0;
    int64_t n_items = 0;
};

}  // namespace

static StructPlan resolve_plan(pripps_engine *e, const pripps_tuning *tune, int precision,
                               const Prepared &p) {
    StructPlan sp;
    const bool expl = tune && (tune->q_per_lane || tune->block || tune->span || tune->variant);
    pripps_tuning_entry te{};
    bool have = false;
    if (!expl) {
        std::lock_guard<std::mutex> lk(e->tune_mu);
        const int la = log2_atoms_of(p.n), ln = log2_nq_of(p.nq);
        for (const auto &x : e->tune_table)
            if (x.precision == precision && x.uniform == (p.uniform ? 1 : 0) && x.log2_atoms == la &&
                x.log2_nq == ln) {
                te = x;
                have = true;
                break;
            }
    }
    int Q = expl ? valid_q(precision, p.uniform, tune->q_per_lane)
                 : have ? valid_q(precision, p.uniform, te.q_per_lane) : 0;
    if (!Q) Q = default_q(precision, p.uniform, p.nq, p.n);
    sp.Q = Q;
    if (Q == kChebQ) cheb_strips(p, sp, precision == PRIPPS_FP64);
    const int base = default_span(p, Q, e->n_cu, Q == kChebQ ? sp.n_units(p.nq) : 0);
    if (expl && tune->span > 0) sp.span = std::max(64, tune->span);
    else if (have) sp.span = shifted_span(base, te.span_shift);
    else sp.span = base;
    // Chebyshev plans size their units themselves unless a span is given
    // (0 in the build statistics)
    sp.span_set = (expl && tune->span > 0) || (have && (Q != kChebQ || te.span_shift != 0));
    if (Q == kChebQ && !sp.span_set) sp.span = 0;
    const int block = expl && tune->block > 0 ? tune->block
                      : have && te.block > 0  ? te.block
                      : Q == kChebQ           ? kChebBlock
                                              : kDefaultBlock;
    sp.block = std::max(8, std::min(block, 4096));
    sp.variant = valid_variant(precision, p.uniform, Q, expl ? tune->variant : (have ? te.variant : 3));
    return sp;
}

// Units of a Chebyshev launch.  A structure's work for one strip group is the
// sequence of its tiles' j ranges in order, each j costing one pass (two for
// hydrated j, whose chunks run the w_j pass too); it is cut into units of
// equal cost -- contiguous runs, so a unit holds one or a few items -- and
// each unit sums into its own partial slots, which the reduction adds in a
// fixed order.  A structure that can fill the GPU gets one unit per wave
// slot of the launch, so that the waves finish together; smaller ones get
// units of about kChebUnitMin j atoms; a span from the caller's tuning sets
// the unit length instead (in j atoms of a two-strip group).  The cut
// depends on the structure alone, so its profiles are bitwise the same alone
// or in any batch.  The queue runs the units costliest first (a stable sort).
constexpr int64_t kChebUnitMin = 48;

// Unit weights by wave age.  A SIMD issues the oldest ready wave first, so of
// the waves that share a SIMD -- one from each of the `ranks` blocks resident
// on a CU -- the oldest runs almost as if alone and the youngest gets what is
// left: with equal units at 3 blocks per CU they ended at about 117, 176 and
// 246 us of a 0.27 ms launch (solvent10k, FP32).  The dispatcher places the
// first n_cu blocks one per CU, then the next n_cu, so block b's age rank on
// its CU is b / n_cu (97.5 % of the waves, measured), and wave slot g runs
// unit g first.  A structure that fills the launch is therefore cut into
// units in the repeating weight pattern of its rank count: the costliest-
// first order puts the heaviest third on the oldest blocks.
// Weights measured best on solvent10k (FP32, P = 26: 0.2523 ms against
// 0.2655 equal, 0.2701 at 2.2 : 1.1 : 0.35; FP64, P = 36: 0.681-0.684 ms at
// 1.8 : 0.4 against 0.693 at 1.6 : 0.5, 0.706 at 2.0 : 0.35), and on
// helix60k (P = 40: 51.2 ms at 2.2 : 1.1 : 0.35, 52.9 at 1.6 : 1.0 : 0.5,
// 57.0 equal).
static double cheb_rank_weight(int ranks, int P, int64_t k) {
    static constexpr double w3[3] = {1.6, 1.0, 0.5};
    static constexpr double w3_long[3] = {2.2, 1.1, 0.35};
    static constexpr double w2[2] = {1.8, 0.4};
    if (ranks == 3) return P >= 36 ? w3_long[k % 3] : w3[k % 3];
    if (ranks == 2) return w2[k % 2];
    return 1.0;
}

// The same cut for two structures: same atoms per class, tiles, strips and
// order (frames of one topology, say).
static bool same_cut(const Prepared &a, const StructPlan &qa, const Prepared &b, const StructPlan &qb) {
    if (a.n != b.n || a.nq != b.nq || a.n_classes != b.n_classes || qa.P != qb.P ||
        qa.span_set != qb.span_set || qa.span != qb.span || qa.group_first != qb.group_first ||
        a.tiles.size() != b.tiles.size())
        return false;
    for (int c = 0; c <= a.n_classes; ++c)
        if (a.cls_start[c] != b.cls_start[c]) return false;
    for (int c = 0; c < a.n_classes; ++c)
        if (a.cls_kind[c] != b.cls_kind[c] || a.cls_wet[c] != b.cls_wet[c]) return false;
    for (size_t t = 0; t < a.tiles.size(); ++t)
        if (a.tiles[t].x != b.tiles[t].x || a.tiles[t].y != b.tiles[t].y || a.tiles[t].z != b.tiles[t].z)
            return false;
    return true;
}

static void plan_cheb_group(LaunchGroup &g, const std::vector<Prepared> &P,
                            const std::vector<StructPlan> &sp, int n_waves, int ranks) {
    struct Pending {
        double cost;
        Unit u;
        int32_t s, strip;
        int32_t cls;  // wave-age class of a rank-weighted unit (0: heaviest), else 0
    };
    std::vector<Pending> pend;
    g.items.clear();
    int prev_s = -1;
    size_t prev_items0 = 0, prev_items1 = 0, prev_pend0 = 0, prev_pend1 = 0;
    for (int s : g.members) {
        const Prepared &p = P[s];
        const StructPlan &q = sp[s];
        const int ng = q.n_units(p.nq);
        if (ng == 0 || p.tiles.empty()) continue;
        if (prev_s >= 0 && same_cut(P[prev_s], sp[prev_s], p, q)) {
            // copy the previous structure's items and units
            const int32_t shift = (int32_t)(g.items.size() - prev_items0);
            for (size_t i = prev_items0; i < prev_items1; ++i) {
                Item it = g.items[i];
                it.s = s;
                g.items.push_back(it);
            }
            for (size_t x = prev_pend0; x < prev_pend1; ++x) {
                Pending pd = pend[x];
                pd.s = s;
                pd.u.item0 += shift;
                pd.u.item1 += shift;
                pend.push_back(pd);
            }
            prev_items0 = prev_items1;
            prev_items1 = g.items.size();
            prev_pend0 = prev_pend1;
            prev_pend1 = pend.size();
            prev_s = s;
            continue;
        }
        const size_t items0 = g.items.size(), pend0 = pend.size();
        // the class segments from j on, each j costing 1 (2 if hydrated)
        auto cost = [&](int32_t a, int32_t b) {
            int64_t c = b - a;
            for (int k = 0; k < p.n_classes; ++k)
                if (p.cls_wet[k]) {
                    const int32_t lo = std::max(a, p.cls_start[k]), hi = std::min(b, p.cls_start[k + 1]);
                    if (hi > lo) c += hi - lo;
                }
            return c;
        };
        // the first j1 > a with cost(a, j1) >= need (need > 0, below cost(a, n))
        auto advance = [&](int32_t a, int64_t need) {
            for (int k = 0; k < p.n_classes; ++k) {
                const int32_t lo = std::max(a, p.cls_start[k]), hi = p.cls_start[k + 1];
                if (hi <= lo) continue;
                const int64_t w = p.cls_wet[k] ? 2 : 1;
                if ((hi - lo) * w >= need) return (int32_t)(lo + (need + w - 1) / w);
                need -= (hi - lo) * w;
            }
            return p.n;
        };
        int64_t W = 0;
        for (const int4 &t : p.tiles) W += cost(t.x, p.n);
        if (W <= 0) continue;
        std::vector<double> f(ng);
        double fs = 0.0;
        for (int u = 0; u < ng; ++u) {
            f[u] = g.fp64 ? 41.0 + 2.0 * q.P + q.unit_strips(u) * (42.0 + q.P)
                          : 30.0 + 2.0 * q.P + q.unit_strips(u) * (29.0 + q.P);
            fs += f[u];
        }
        const double fref = g.fp64 ? 41.0 + 2.0 * q.P + 2.0 * (42.0 + q.P)
                                   : 30.0 + 2.0 * q.P + 2.0 * (29.0 + q.P);
        const double wref = (double)W * fs / fref;  // cost in j atoms of a two-strip group
        const int64_t us =
            q.span_set ? std::max<int64_t>(1, (int64_t)std::ceil(wref / q.span))
                       : std::max<int64_t>(1, std::min<int64_t>(n_waves, (int64_t)(wref / (double)kChebUnitMin)));
        g.unit_len = std::max(g.unit_len, (int64_t)std::ceil(wref / (double)us));
        // a structure that fills the launch (one unit per wave slot) is cut
        // into units weighted by wave age, classes in turn over all its
        // units, so that each class holds a third (for three ranks) of them
        const bool by_rank = !q.span_set && us == n_waves && (ranks == 2 || ranks == 3);
        int64_t kk = 0;  // the structure's unit count so far
        for (int u = 0; u < ng; ++u) {
            const int nk = q.unit_strips(u);
            const int strip = q.unit_first(u);
            int64_t ug = std::max<int64_t>(1, std::llround((double)us * f[u] / fs));
            ug = std::min(ug, W);
            std::vector<double> cw((size_t)ug);
            double acc_w = 0.0;
            for (int64_t k = 0; k < ug; ++k) {
                acc_w += by_rank ? cheb_rank_weight(ranks, g.P, kk + k) : 1.0;
                cw[(size_t)k] = acc_w;
            }
            size_t t = 0;
            int32_t j = p.tiles[0].x;
            int64_t rest = cost(j, p.n);  // cost of the current tile from j on
            int64_t done = 0, prev = 0;
            for (int64_t k = 0; k < ug && t < p.tiles.size(); ++k) {
                const int64_t target =
                    k + 1 == ug ? W : std::min<int64_t>(W, (int64_t)((double)W * cw[(size_t)k] / acc_w));
                const int32_t item0 = (int32_t)g.items.size();
                while (done < target && t < p.tiles.size()) {
                    const int4 &tl = p.tiles[t];
                    int32_t j1 = p.n;
                    int64_t c = rest;
                    if (done + rest > target) {
                        j1 = advance(j, target - done);
                        c = cost(j, j1);
                    }
                    g.items.push_back(Item{s, tl.x, tl.y, tl.z, j, j1, strip, nk - 1});
                    done += c;
                    rest -= c;
                    if (j1 == p.n) {
                        if (++t < p.tiles.size()) {
                            j = p.tiles[t].x;
                            rest = cost(j, p.n);
                        }
                    } else {
                        j = j1;
                    }
                }
                const int32_t item1 = (int32_t)g.items.size();
                if (item1 > item0)
                    pend.push_back(Pending{(double)(done - prev) * f[u], Unit{item0, item1, 0, nk}, s, strip,
                                           by_rank ? (int32_t)((kk + k) % ranks) : 0});
                prev = done;
            }
            kk += ug;
        }
        prev_s = s;
        prev_items0 = items0;
        prev_items1 = g.items.size();
        prev_pend0 = pend0;
        prev_pend1 = pend.size();
    }
    // Queue order: costliest first, ties in generation order; with rank-
    // weighted units, per class c (heaviest first) its costliest n_waves /
    // ranks units first -- the units that wave slots [c, c + 1) n_waves /
    // ranks, the blocks of age rank c, take without the queue -- and then
    // every other unit costliest first.
    bool weighted = false;
    for (const Pending &pd : pend) weighted |= pd.cls != 0;
    std::vector<int32_t> order;
    order.reserve(pend.size());
    {
        std::vector<std::pair<double, int32_t>> byc(pend.size());
        for (size_t x = 0; x < pend.size(); ++x) byc[x] = {-pend[x].cost, (int32_t)x};
        std::sort(byc.begin(), byc.end());
        if (weighted) {
            const int64_t per = n_waves / ranks;
            std::vector<char> first(pend.size(), 0);
            for (int c = 0; c < ranks; ++c) {
                int64_t taken = 0;
                for (const auto &e : byc)
                    if (taken < per && pend[e.second].cls == c) {
                        order.push_back(e.second);
                        first[e.second] = 1;
                        ++taken;
                    }
            }
            for (const auto &e : byc)
                if (!first[e.second]) order.push_back(e.second);
        } else {
            for (const auto &e : byc) order.push_back(e.second);
        }
    }
    // Records: one per class run of each item, numbered unit by unit in
    // generation order (a unit's records are consecutive from its rec0, in
    // the order in which the kernel writes them).  Combos (structure, strip
    // group, i kind, j kind), numbered in that order, list their records in
    // record order (a counting sort).  Neither depends on the queue order or
    // on the batch, so a structure's sums are the same alone or in any batch.
    g.RS = ((g.fp64 ? 1 : 2) * g.P + 63) / 64 * 64;
    const int nm = (int)g.members.size();
    std::vector<int> local(P.size(), -1);
    std::vector<int64_t> cbase(nm + 1, 0);
    for (int k = 0; k < nm; ++k) {
        const int s = g.members[k];
        local[s] = k;
        cbase[k + 1] = cbase[k] + (int64_t)sp[s].n_units(P[s].nq) * P[s].K * P[s].K;
    }
    // strip -> group index of each member (strip groups are contiguous)
    std::vector<std::vector<int32_t>> group_of(nm);
    for (int k = 0; k < nm; ++k) {
        const StructPlan &q = sp[g.members[k]];
        const int ns = q.n_strips(P[g.members[k]].nq);
        group_of[k].assign(std::max(ns, 1), 0);
        for (int u = 0; u < q.n_units(P[g.members[k]].nq); ++u)
            for (int st = q.unit_first(u); st < q.unit_first(u) + q.unit_strips(u); ++st) group_of[k][st] = u;
    }
    std::vector<int32_t> rec_combo;
    for (Pending &pd : pend) {
        Unit &u = pd.u;
        u.rec0 = (int32_t)rec_combo.size();
        for (int32_t i = u.item0; i < u.item1; ++i) {
            const Item &it = g.items[i];
            const Prepared &p = P[it.s];
            const int k = local[it.s];
            const int64_t base = cbase[k] + ((int64_t)group_of[k][it.strip] * p.K + it.ikind) * p.K;
            for (int c = 0; c < p.n_classes; ++c) {
                const int32_t lo = std::max(it.j0, p.cls_start[c]);
                const int32_t hi = std::min(it.j1, p.cls_start[c + 1]);
                if (lo < hi) rec_combo.push_back((int32_t)(base + p.cls_kind[c]));
            }
        }
    }
    g.n_records = (int64_t)rec_combo.size();
    g.units.resize(pend.size());
    for (size_t x = 0; x < order.size(); ++x) g.units[x] = pend[order[x]].u;
    const int64_t n_combos = cbase[nm];
    std::vector<int32_t> cptr(n_combos + 1, 0);
    for (int32_t c : rec_combo) cptr[c + 1]++;
    for (int64_t c = 0; c < n_combos; ++c) cptr[c + 1] += cptr[c];
    std::vector<int32_t> crec(rec_combo.size());
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (size_t r = 0; r < rec_combo.size(); ++r) crec[fill[rec_combo[r]]++] = (int32_t)r;
    }
    // stage 1 chunks, combo by combo; stage 2 blocks per strip and non-empty
    // combo of its group; stage 3 blocks per strip (every strip, with or
    // without work)
    g.chunk_recs.assign(crec.begin(), crec.end());
    g.chunk_first.assign(1, 0);
    std::vector<int32_t> c_first(n_combos + 1, 0);
    for (int64_t c = 0; c < n_combos; ++c) {
        c_first[c] = (int32_t)g.chunk_first.size() - 1;
        for (int32_t x = cptr[c]; x < cptr[c + 1]; x += kFinChunk)
            g.chunk_first.push_back(std::min(cptr[c + 1], x + kFinChunk));
    }
    c_first[n_combos] = (int32_t)g.chunk_first.size() - 1;
    g.sb_a.clear();
    g.sb_b.clear();
    g.st.clear();
    for (int k = 0; k < nm; ++k) {
        const int s = g.members[k];
        const StructPlan &q = sp[s];
        const int K = P[s].K;
        const int ng = q.n_units(P[s].nq);
        for (int u = 0; u < ng; ++u) {
            const int gf = q.unit_first(u);
            for (int ks = 0; ks < q.unit_strips(u); ++ks) {
                const int32_t b0 = (int32_t)g.sb_a.size();
                for (int ik = 0; ik < K; ++ik)
                    for (int jk = 0; jk < K; ++jk) {
                        const int64_t c = cbase[k] + ((int64_t)u * K + ik) * K + jk;
                        if (cptr[c + 1] == cptr[c]) continue;
                        g.sb_a.push_back(int4{s, gf + ks, ks, ik});
                        g.sb_b.push_back(int4{jk, c_first[c], c_first[c + 1], 0});
                    }
                g.st.push_back(int4{s, gf + ks, b0, (int32_t)g.sb_a.size()});
            }
        }
    }
}

static void plan_group(LaunchGroup &g, const std::vector<Prepared> &P,
                       const std::vector<StructPlan> &sp) {
    // Queue order: largest span first, and within one span length the order
    // of generation (a stable counting sort: a structure's items keep one
    // relative order in any batch).  The items are generated twice, once to
    // count them per span length and once to write each straight to its
    // place, so no unsorted copy of the list is made.
    auto each_item = [&](auto &&emit) {
        for (int s : g.members) {
            const Prepared &p = P[s];
            const int nu = sp[s].n_units(p.nq);
            const int span = sp[s].span;
            for (const int4 &t : p.tiles)
                for (int32_t j0 = t.x; j0 < p.n; j0 += span)
                    emit(s, t, j0, std::min<int32_t>(p.n, j0 + span), nu);
        }
    };
    int maxlen = 0;
    for (int s : g.members) maxlen = std::max(maxlen, sp[s].span);
    std::vector<int64_t> pos(maxlen + 2, 0);
    each_item([&](int, const int4 &, int32_t j0, int32_t j1, int nu) {
        pos[maxlen - (j1 - j0) + 1] += nu;
    });
    for (int x = 0; x <= maxlen; ++x) pos[x + 1] += pos[x];
    g.items.resize((size_t)pos[maxlen + 1]);
    Item *items = g.items.data();
    each_item([&](int s, const int4 &t, int32_t j0, int32_t j1, int nu) {
        int64_t &w = pos[maxlen - (j1 - j0)];
        for (int u = 0; u < nu; ++u) items[w++] = Item{s, t.x, t.y, t.z, j0, j1, u, 0};
    });
    g.n_slots = (int64_t)g.items.size();  // item w sums into slot w
    // CSR of items per (member, strip), in queue order
    const int nm = (int)g.members.size();
    std::vector<int64_t> base(nm + 1, 0);
    std::vector<int> local(P.size(), -1);
    for (int k = 0; k < nm; ++k) {
        local[g.members[k]] = k;
        base[k + 1] = base[k] + sp[g.members[k]].n_strips(P[g.members[k]].nq);
    }
    const int64_t n_strips = base[nm];
    std::vector<int32_t> cnt(n_strips + 1, 0);
    for (const Item &it : g.items) cnt[base[local[it.s]] + it.strip + 1]++;
    for (int64_t x = 0; x < n_strips; ++x) cnt[x + 1] += cnt[x];
    g.csr_ptr.assign(cnt.begin(), cnt.end());
    g.csr_items.assign(g.items.size(), 0);
    std::vector<int32_t> fillp(cnt.begin(), cnt.end() - 1);
    for (size_t w = 0; w < g.items.size(); ++w)
        g.csr_items[fillp[base[local[g.items[w].s]] + g.items[w].strip]++] = (int32_t)w;
    g.stage_ptr.assign(n_strips + 1, 0);
    for (int64_t x = 0; x < n_strips; ++x) {
        const int nch = (g.csr_ptr[x + 1] - g.csr_ptr[x] + kFinChunk - 1) / kFinChunk;
        g.stage_ptr[x + 1] = g.stage_ptr[x] + nch;
    }
    g.strip_struct.resize(n_strips);
    g.strip_index.resize(n_strips);
    for (int k = 0; k < nm; ++k)
        for (int64_t st = 0; st < base[k + 1] - base[k]; ++st) {
            g.strip_struct[base[k] + st] = g.members[k];
            g.strip_index[base[k] + st] = (int32_t)st;
        }
}

// Plan and lay out one submission.  user_bytes reserves a caller region at
// the end of the uploaded part (B.o_user).
static void plan_build(pripps_engine *e, const std::vector<Prepared> &P, int precision,
                       const pripps_tuning *tune, size_t user_bytes, BuildPlan &B) {
    const int ns = (int)P.size();
    B.ns = ns;
    B.fp64 = precision != PRIPPS_FP32_COMPENSATED;
    B.plain = precision == PRIPPS_FP64_PLAIN;
    B.sp.assign(ns, StructPlan{});
    if (!B.plain) {
        for (int s = 0; s < ns; ++s) {
            // consecutive structures with the same plan inputs (frames of one
            // topology) take the plan of the one before
            const Prepared &a = P[s];
            const Prepared *b = s > 0 ? &P[s - 1] : nullptr;
            if (b && a.n == b->n && a.nq == b->nq && a.uniform == b->uniform && a.rmax == b->rmax &&
                a.tiles.size() == b->tiles.size() &&
                std::memcmp(a.q.data(), b->q.data(), (size_t)a.nq * sizeof(double)) == 0)
                B.sp[s] = B.sp[s - 1];
            else
                B.sp[s] = resolve_plan(e, tune, precision, P[s]);
            LaunchGroup *g = nullptr;
            for (auto &x : B.groups)
                if (x.Q == B.sp[s].Q && x.uniform == P[s].uniform && x.variant == B.sp[s].variant &&
                    x.P == B.sp[s].P)
                    g = &x;
            if (!g) {
                B.groups.emplace_back();
                g = &B.groups.back();
                g->Q = B.sp[s].Q;
                g->uniform = P[s].uniform;
                g->variant = B.sp[s].variant;
                g->P = B.sp[s].P;
                g->fp64 = B.fp64;
            }
            g->members.push_back(s);
        }
        for (auto &g : B.groups) {
            if (g.Q == kChebQ) {
                int bpc = 1;
                if ((g.fp64 ? cheb64_occupancy(g.P, &bpc) : cheb_occupancy(g.P, &bpc)) != hipSuccess) bpc = 1;
                plan_cheb_group(g, P, B.sp, e->n_cu * 4 * std::max(1, bpc), std::max(1, bpc));
                B.n_items += (int64_t)g.units.size();
            } else {
                plan_group(g, P, B.sp);
                B.n_items += (int64_t)g.items.size();
            }
        }
    }
    Layout L;
    const size_t atom_bytes = B.fp64 ? sizeof(Atom<double>) : sizeof(Atom<float>);
    for (auto *v : {&B.o_atoms, &B.o_ffa, &B.o_ffe, &B.o_ffw, &B.o_kn, &B.o_kw, &B.o_tiles, &B.o_pp,
                    &B.o_sf, &B.o_sg})
        v->assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
        B.o_atoms[s] = L.take(std::max<size_t>(1, (size_t)P[s].n) * atom_bytes);
        B.o_ffa[s] = L.take(P[s].ffa.size() * 8 + 8);
        B.o_ffe[s] = L.take(P[s].ffe.size() * 8 + 8);
        B.o_ffw[s] = L.take(P[s].ffw.size() * 8 + 8);
        B.o_kn[s] = L.take(P[s].K * 8);
        B.o_kw[s] = L.take(P[s].K * 8);
        if (!B.plain && B.sp[s].Q == kChebQ) {
            B.o_sf[s] = L.take(B.sp[s].strip_first.size() * 4);
            B.o_sg[s] = L.take(B.sp[s].strip_geo.size() * 8 + 8);
        }
        if (B.plain) B.o_tiles[s] = L.take(P[s].tiles.size() * sizeof(int4) + 16);
    }
    B.o_structs = L.take(ns * sizeof(DevStruct));
    for (auto &g : B.groups) {
        g.o_items = L.take(g.items.size() * sizeof(Item) + 16);
        g.o_units = L.take(g.units.size() * sizeof(Unit) + 16);
        g.o_chunk_recs = L.take(g.chunk_recs.size() * 4 + 4);
        g.o_chunk_first = L.take(g.chunk_first.size() * 4 + 4);
        g.o_sb_a = L.take(g.sb_a.size() * sizeof(int4) + 16);
        g.o_sb_b = L.take(g.sb_b.size() * sizeof(int4) + 16);
        g.o_st = L.take(g.st.size() * sizeof(int4) + 16);
        g.o_csr_ptr = L.take(g.csr_ptr.size() * 4 + 4);
        g.o_csr_items = L.take(g.csr_items.size() * 4 + 4);
        g.o_ss = L.take(g.strip_struct.size() * 4 + 4);
        g.o_si = L.take(g.strip_index.size() * 4 + 4);
        g.o_sp = L.take(g.stage_ptr.size() * 4 + 4);
        g.o_counter = L.take(256);
    }
    B.o_user = L.take(std::max<size_t>(user_bytes, 8));
    B.upload_bytes = L.off;
    if (B.plain) {
        for (int s = 0; s < ns; ++s)
            B.o_pp[s] = L.take(std::max<size_t>(1, P[s].tiles.size() * (size_t)P[s].nq * kProfiles) * 8);
    }
    for (auto &g : B.groups) {
        g.o_partials = L.take(std::max<int64_t>(1, g.n_slots) * kProfiles * g.Q * 8);
        if (g.Q == kChebQ) {
            const int64_t n_chunks = (int64_t)g.chunk_first.size() - 1;
            g.o_records = L.take(std::max<int64_t>(1, g.n_records) * kChebSets * g.RS * 8);
            g.o_chunk_tot = L.take(std::max<int64_t>(1, n_chunks) * kChebSets * g.RS * 8);
            g.o_sb_out = L.take(std::max<size_t>(1, g.sb_a.size()) * kProfiles * kChebQ * 8);
        }
        const size_t nch = g.stage_ptr.empty() ? 0 : (size_t)g.stage_ptr.back();
        g.o_stage = L.take(std::max<size_t>(1, nch) * kProfiles * g.Q * 8);
    }
    B.total_bytes = L.off;
}

// Fill the pinned staging buffer for plan B: atoms (converted), tables,
// structure records (pointing at the arena, q[s] and prof[s]) and, unless the
// buffer holds them from an earlier staging of the same plan, the work lists.
static void stage_build(const BuildPlan &B, const std::vector<Prepared> &P, char *hbase,
                        char *dbase, const std::vector<const double *> &d_q,
                        const std::vector<double *> &d_prof, bool work_lists = true) {
    std::vector<DevStruct> ds(B.ns);
    for (int s = 0; s < B.ns; ++s) {
        const Prepared &p = P[s];
        if (B.fp64) {
            std::memcpy(hbase + B.o_atoms[s], p.pos.data(), p.pos.size() * 8);
        } else {
            float *f = (float *)(hbase + B.o_atoms[s]);
            for (size_t i = 0; i < p.pos.size(); ++i) f[i] = (float)p.pos[i];
        }
        std::memcpy(hbase + B.o_ffa[s], p.ffa.data(), p.ffa.size() * 8);
        if (!p.ffe.empty()) std::memcpy(hbase + B.o_ffe[s], p.ffe.data(), p.ffe.size() * 8);
        if (!p.ffw.empty()) std::memcpy(hbase + B.o_ffw[s], p.ffw.data(), p.ffw.size() * 8);
        std::memcpy(hbase + B.o_kn[s], p.kind_n.data(), p.K * 8);
        std::memcpy(hbase + B.o_kw[s], p.kind_w.data(), p.K * 8);
        if (B.plain) std::memcpy(hbase + B.o_tiles[s], p.tiles.data(), p.tiles.size() * sizeof(int4));
        const bool cheb = !B.plain && B.sp[s].Q == kChebQ;
        if (cheb) {
            const StructPlan &q = B.sp[s];
            std::memcpy(hbase + B.o_sf[s], q.strip_first.data(), q.strip_first.size() * 4);
            std::memcpy(hbase + B.o_sg[s], q.strip_geo.data(), q.strip_geo.size() * 8);
        }
        DevStruct &d = ds[s];
        std::memset(&d, 0, sizeof(d));
        d.atoms = dbase + B.o_atoms[s];
        d.q = d_q[s];
        d.ffa = (const double *)(dbase + B.o_ffa[s]);
        d.ffe = p.has_e ? (const double *)(dbase + B.o_ffe[s]) : nullptr;
        d.ffw = p.has_h ? (const double *)(dbase + B.o_ffw[s]) : nullptr;
        d.profiles = d_prof[s];
        d.kind_n = (const double *)(dbase + B.o_kn[s]);
        d.kind_w = (const double *)(dbase + B.o_kw[s]);
        d.sum_w2 = p.sum_w2;
        d.q0 = p.q0;
        d.dq = p.dq;
        d.n = p.n;
        d.nq = p.nq;
        d.K = p.K;
        d.uniform = p.uniform ? 1 : 0;
        d.small_h = p.small_h ? 1 : 0;
        d.sin_direct = p.sin_direct ? 1 : 0;
        d.n_classes = p.n_classes;
        d.n_strips = B.plain ? 0 : B.sp[s].n_strips(p.nq);
        d.strip_first = cheb ? (const int32_t *)(dbase + B.o_sf[s]) : nullptr;
        d.strip_geo = cheb ? (const double *)(dbase + B.o_sg[s]) : nullptr;
        d.block = B.sp[s].block;
        for (int c = 0; c <= p.n_classes; ++c) d.cls_start[c] = p.cls_start[c];
        for (int c = p.n_classes + 1; c <= kMaxClasses; ++c) d.cls_start[c] = p.n;
        for (int c = 0; c < p.n_classes; ++c) {
            d.cls_kind[c] = p.cls_kind[c];
            d.cls_wet[c] = p.cls_wet[c];
        }
    }
    std::memcpy(hbase + B.o_structs, ds.data(), B.ns * sizeof(DevStruct));
    if (!work_lists) return;
    auto put = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) std::memcpy(hbase + off, src, bytes);
    };
    for (auto &g : B.groups) {
        std::memset(hbase + g.o_counter, 0, 256);  // queue counters start at zero
        put(g.o_items, g.items.data(), g.items.size() * sizeof(Item));
        put(g.o_units, g.units.data(), g.units.size() * sizeof(Unit));
        put(g.o_chunk_recs, g.chunk_recs.data(), g.chunk_recs.size() * 4);
        put(g.o_chunk_first, g.chunk_first.data(), g.chunk_first.size() * 4);
        put(g.o_sb_a, g.sb_a.data(), g.sb_a.size() * sizeof(int4));
        put(g.o_sb_b, g.sb_b.data(), g.sb_b.size() * sizeof(int4));
        put(g.o_st, g.st.data(), g.st.size() * sizeof(int4));
        put(g.o_csr_ptr, g.csr_ptr.data(), g.csr_ptr.size() * 4);
        put(g.o_csr_items, g.csr_items.data(), g.csr_items.size() * 4);
        put(g.o_ss, g.strip_struct.data(), g.strip_struct.size() * 4);
        put(g.o_si, g.strip_index.data(), g.strip_index.size() * 4);
        put(g.o_sp, g.stage_ptr.data(), g.stage_ptr.size() * 4);
    }
}

static int make_launches(pripps_engine *e, const BuildPlan &B, char *dbase,
                         std::vector<LaunchRecord> &launches) {
    launches.clear();
    for (auto &g : B.groups) {
        LaunchRecord r;
        r.fp64 = B.fp64;
        r.Q = g.Q;
        r.P = g.P;
        r.uniform = g.uniform;
        r.counter = (int *)(dbase + g.o_counter);
        const bool cheb = g.Q == kChebQ;
        r.n_items = cheb ? (int64_t)g.units.size() : (int64_t)g.items.size();
        r.n_strips = (int)g.strip_struct.size();
        TiledArgs &a = r.targs;
        a.structs = (const DevStruct *)(dbase + B.o_structs);
        a.items = (const Item *)(dbase + g.o_items);
        a.n_items = r.n_items;
        a.counter = r.counter;
        a.partials = (double *)(dbase + g.o_partials);
        a.variant = g.variant;
        a.units = cheb ? (const Unit *)(dbase + g.o_units) : nullptr;
        if (cheb) {
            a.partials = (double *)(dbase + g.o_records);
            ChebReduceArgs &c = r.cargs;
            c.structs = a.structs;
            c.records = a.partials;
            c.RS = g.RS;
            c.P = g.P;
            c.chunk_recs = (const int32_t *)(dbase + g.o_chunk_recs);
            c.chunk_first = (const int32_t *)(dbase + g.o_chunk_first);
            c.n_chunks = (int)g.chunk_first.size() - 1;
            c.chunk_tot = (double *)(dbase + g.o_chunk_tot);
            c.sb_a = (const int4 *)(dbase + g.o_sb_a);
            c.sb_b = (const int4 *)(dbase + g.o_sb_b);
            c.n_sb = (int)g.sb_a.size();
            c.sb_out = (double *)(dbase + g.o_sb_out);
            c.st = (const int4 *)(dbase + g.o_st);
            c.n_st = (int)g.st.size();
        }
        int bpc = 1;
        if (g.Q == kChebQ) HIPCHK(g.fp64 ? cheb64_occupancy(g.P, &bpc) : cheb_occupancy(g.P, &bpc));
        else HIPCHK(tiled_occupancy(B.fp64, g.Q, g.uniform, g.variant, &bpc));
        bpc = std::max(1, bpc);
        const int64_t want = (r.n_items + 3) / 4;
        r.blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)e->n_cu * bpc));
        FinalizeArgs &f = r.fargs;
        f.structs = a.structs;
        f.csr_ptr = (const int32_t *)(dbase + g.o_csr_ptr);
        f.csr_items = (const int32_t *)(dbase + g.o_csr_items);
        f.partials = a.partials;
        f.strip_struct = (const int32_t *)(dbase + g.o_ss);
        f.strip_index = (const int32_t *)(dbase + g.o_si);
        f.stage_ptr = (const int32_t *)(dbase + g.o_sp);
        f.stage = (double *)(dbase + g.o_stage);
        f.n_strips = g.stage_ptr.empty() ? 0 : (int)(g.stage_ptr.size() - 1);
        f.total_chunks = g.stage_ptr.empty() ? 0 : g.stage_ptr.back();
        f.Q = g.Q;
        launches.push_back(r);
    }
    return PRIPPS_OK;
}

// The kernels of one submission (after its upload), on `st`.
static hipError_t enqueue_kernels(const BuildPlan &B, const std::vector<Prepared> &P, char *dbase,
                                  const std::vector<LaunchRecord> &launches, hipStream_t st) {
    hipError_t err = hipSuccess;
    if (B.plain) {
        for (int s = 0; s < B.ns && err == hipSuccess; ++s) {
            if (P[s].nq == 0) continue;
            PlainArgs a;
            a.sd = (const DevStruct *)(dbase + B.o_structs) + s;
            a.tiles = (const int4 *)(dbase + B.o_tiles[s]);
            a.n_tiles = (int)P[s].tiles.size();
            a.partials = (double *)(dbase + B.o_pp[s]);
            err = launch_plain(a, P[s].nq, st);
        }
    } else {
        for (auto &r : launches) {
            if (err == hipSuccess) err = run_launch(r, st, nullptr, nullptr);
        }
    }
    return err;
}

static int build_models(pripps_engine *e, int32_t ns, const pripps_debye_input *inputs,
                        int32_t precision, const pripps_tuning *tune, pripps_debye_model **out) {
    const auto t_start = std::chrono::steady_clock::now();
    if (!e) return fail(PRIPPS_ERR_USAGE, "engine is NULL");
    if (!out) return fail(PRIPPS_ERR_USAGE, "out is NULL");
    if (ns < 0) return fail(PRIPPS_ERR_USAGE, "negative structure count");
    if (precision != PRIPPS_FP32_COMPENSATED && precision != PRIPPS_FP64 &&
        precision != PRIPPS_FP64_PLAIN)
        return fail(PRIPPS_ERR_USAGE, "unknown precision " + std::to_string(precision));
    for (int s = 0; s < ns; ++s) out[s] = nullptr;
    if (ns == 0) return PRIPPS_OK;

    // host preparation, in parallel over structures for a batch (prepare()
    // is a pure function of its input; errors are reported for the first
    // failing structure in input order)
    std::vector<Prepared> P(ns);
    {
        const int nt = std::max(1, std::min(ns / 2, std::min(8, (int)std::thread::hardware_concurrency())));
        std::vector<int> rcs(ns, PRIPPS_OK);
        std::vector<std::string> msgs(ns);
        auto part = [&](int t) {
            for (int s = t; s < ns; s += nt) {
                rcs[s] = prepare(&inputs[s], P[s]);
                if (rcs[s]) msgs[s] = g_last_error;  // this thread's message
            }
        };
        if (nt == 1) {
            part(0);
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < nt; ++t) pool.emplace_back(part, t);
            for (auto &th : pool) th.join();
        }
        for (int s = 0; s < ns; ++s)
            if (rcs[s]) {
                g_last_error = ns > 1 ? "structure " + std::to_string(s) + ": " + msgs[s] : msgs[s];
                return rcs[s];
            }
    }

    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));

    // models (persistent device memory)
    std::vector<pripps_debye_model *> models(ns, nullptr);
    // a failed build's blocks go back to the pool only after the stream has
    // drained (nothing in flight may still write them)
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(e->stream);
        for (auto *m : models) {
            if (m) {
                model_release_block(m);
                delete m;
            }
        }
    };
    for (int s = 0; s < ns; ++s) {
        auto *m = new pripps_debye_model();
        models[s] = m;
        m->engine = e;
        m->nq = P[s].nq;
        m->q = P[s].q;
        hipError_t err = model_alloc(m);
        if (err != hipSuccess) {
            cleanup();
            return hip_fail(err, "model buffer");
        }
    }

    BuildPlan B;
    plan_build(e, P, precision, tune, 0, B);
    int rc = ensure_arena(e, B.total_bytes);
    if (!rc) rc = ensure_staging(e, B.upload_bytes);
    if (rc) {
        cleanup();
        return rc;
    }
    char *dbase = (char *)e->arena;
    char *hbase = (char *)e->staging;
    std::vector<const double *> dq(ns);
    std::vector<double *> dprof(ns);
    for (int s = 0; s < ns; ++s) {
        dq[s] = models[s]->d_q();
        dprof[s] = models[s]->d_prof();
    }
    stage_build(B, P, hbase, dbase, dq, dprof);
    std::vector<LaunchRecord> launches;
    rc = make_launches(e, B, dbase, launches);
    if (rc) {
        cleanup();
        return rc;
    }

    auto run = [&]() -> int {
        hipStream_t st = e->stream;
        HIPCHK(hipMemcpyAsync(dbase, hbase, B.upload_bytes, hipMemcpyHostToDevice, st));
        for (int s = 0; s < ns; ++s) {
            if (P[s].nq > 0)
                HIPCHK(hipMemcpyAsync(models[s]->d_q(), P[s].q.data(), P[s].nq * 8,
                                      hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipEventRecord(e->ev0, st));
        HIPCHK(enqueue_kernels(B, P, dbase, launches, st));
        HIPCHK(hipEventRecord(e->ev1, st));
        HIPCHK(hipStreamSynchronize(st));
        return PRIPPS_OK;
    };
    e->build_serial++;
    e->last_plain = B.plain;
    e->last_launches = launches;
    rc = run();
    if (rc) {
        e->build_serial++;  // nothing valid to replay
        e->last_launches.clear();
        cleanup();
        return rc;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
    const double total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    for (int s = 0; s < ns; ++s) {
        e->refs.fetch_add(1);
        models[s]->build_serial = e->build_serial;
        pripps_build_stats &S = models[s]->stats;
        S.kernel_ms = ms;
        S.total_ms = total_ms;
        S.pair_q_terms = P[s].pair_q_terms;
        S.q_per_lane = B.plain ? 0 : B.sp[s].Q;
        S.block = B.plain ? 0 : B.sp[s].block;
        S.span = B.plain ? 0 : B.sp[s].span;
        S.cheb_order = B.plain ? 0 : B.sp[s].P;
        S.work_items = B.n_items;
        out[s] = models[s];
    }
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_build(pripps_engine *e, const pripps_debye_input *input,
                                        int32_t precision, const pripps_tuning *tuning,
                                        pripps_debye_model **out) {
    if (!out) return fail(PRIPPS_ERR_USAGE, "out is NULL");
    *out = nullptr;
    if (!input) return fail(PRIPPS_ERR_USAGE, "input is NULL");
    return build_models(e, 1, input, precision, tuning, out);
}

extern "C" int pripps_debye_model_build_batch(pripps_engine *e, int32_t n,
                                              const pripps_debye_input *inputs, int32_t precision,
                                              const pripps_tuning *tuning,
                                              pripps_debye_model **out) {
    if (n > 0 && !inputs) return fail(PRIPPS_ERR_USAGE, "inputs is NULL");
    if (precision == PRIPPS_FP64_PLAIN) {
        // the plain path is single-structure by design: one launch per structure
        for (int32_t s = 0; s < n; ++s) {
            int rc = build_models(e, 1, inputs + s, precision, tuning, out + s);
            if (rc) {
                for (int32_t t = 0; t < s; ++t) {
                    pripps_debye_model_free(out[t]);
                    out[t] = nullptr;
                }
                return rc;
            }
        }
        return PRIPPS_OK;
    }
    return build_models(e, n, inputs, precision, tuning, out);
}

// ---------------------------------------------------------------------------
// build without device allocation
// ---------------------------------------------------------------------------
// Every device buffer of pripps_debye_build_noalloc lives in this array:
// atoms, tables, structure record, work lists, queue head, item partials,
// chunk totals, q grid and the output profiles.  Each GPU the library runs
// on holds its own copy for the life of the process.  In host code the name
// `scratch` is only a host-side stand-in; kernels receive the device address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[PRIPPS_NOALLOC_SCRATCH_BYTES];

namespace {
struct NoAllocDevice {
    std::mutex mu;          // one call at a time per GPU (the array is shared)
    char *ptr = nullptr;    // device address of `scratch` on this GPU
};

NoAllocDevice &noalloc_device(int device) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<NoAllocDevice>> devs;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)devs.size() <= device) devs.resize((size_t)device + 1);
    if (!devs[device]) devs[device].reset(new NoAllocDevice());
    return *devs[device];
}
}  // namespace

extern "C" int pripps_debye_build_noalloc(pripps_engine *e, const pripps_debye_input *input,
                                          int32_t precision, const pripps_tuning *tune,
                                          double *profiles_out) {
    if (!e || !input) return fail(PRIPPS_ERR_USAGE, "engine and input must be non-NULL");
    if (precision != PRIPPS_FP32_COMPENSATED && precision != PRIPPS_FP64)
        return fail(PRIPPS_ERR_USAGE, "noalloc build: precision must be FP32_COMPENSATED or FP64");
    std::vector<Prepared> P(1);
    int rc = prepare(input, P[0]);
    if (rc) return rc;
    const int32_t nq = P[0].nq;
    if (nq > 0 && !profiles_out) return fail(PRIPPS_ERR_USAGE, "profiles_out is NULL");
    if (nq == 0) return PRIPPS_OK;
    // the q grid and the profiles sit in the caller region of the upload
    const size_t user = (size_t)7 * nq * sizeof(double);
    BuildPlan B;
    plan_build(e, P, precision, tune, user, B);
    // Without an explicit span, a plan whose partials do not fit takes longer
    // j-spans (fewer items, so fewer partial slots) until it fits; a
    // Chebyshev plan longer units (fewer coefficient records), up to one
    // unit per strip group.
    const bool span_given = tune && tune->span > 0;
    auto cur_span = [&](const BuildPlan &b) -> int64_t {
        if (b.sp[0].Q != kChebQ) return b.sp[0].span;
        return b.groups.empty() ? INT32_MAX : std::max<int64_t>(b.sp[0].span, b.groups[0].unit_len);
    };
    const int64_t max_span = B.sp[0].Q == kChebQ ? INT32_MAX / 2 : 16384;
    while (B.total_bytes > PRIPPS_NOALLOC_SCRATCH_BYTES && !span_given && cur_span(B) < max_span) {
        pripps_tuning t{};
        t.q_per_lane = B.sp[0].Q;
        t.block = B.sp[0].block;
        t.span = (int32_t)std::min<int64_t>(max_span, 2 * std::max<int64_t>(64, cur_span(B)));
        t.variant = B.sp[0].variant;
        BuildPlan wider;
        plan_build(e, P, precision, &t, user, wider);
        B = std::move(wider);
    }
    if (B.total_bytes > PRIPPS_NOALLOC_SCRATCH_BYTES)
        return fail(PRIPPS_ERR_OUT_OF_MEMORY,
                    "noalloc build: this structure needs " + std::to_string(B.total_bytes) +
                        " bytes of device scratch; the fixed scratch holds " +
                        std::to_string((size_t)PRIPPS_NOALLOC_SCRATCH_BYTES) +
                        " (use pripps_debye_model_build)");
    // host staging of the upload (pageable: no pinned allocation)
    std::vector<char> host(B.upload_bytes);

    NoAllocDevice &D = noalloc_device(e->device);
    std::lock_guard<std::mutex> dlock(D.mu);
    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));
    if (!D.ptr) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PRIPPS_ERR_DEVICE, "noalloc build: no device address for the scratch array");
        D.ptr = (char *)p;
    }
    char *dbase = D.ptr;
    double *d_q = (double *)(dbase + B.o_user);
    double *d_prof = d_q + nq;
    std::vector<const double *> dq(1, d_q);
    std::vector<double *> dprof(1, d_prof);
    stage_build(B, P, host.data(), dbase, dq, dprof);
    std::memcpy(host.data() + B.o_user, P[0].q.data(), (size_t)nq * sizeof(double));
    std::vector<LaunchRecord> launches;
    rc = make_launches(e, B, dbase, launches);
    if (rc) return rc;
    hipStream_t st = e->stream;
    HIPCHK(hipMemcpyAsync(dbase, host.data(), B.upload_bytes, hipMemcpyHostToDevice, st));
    HIPCHK(enqueue_kernels(B, P, dbase, launches, st));
    HIPCHK(hipMemcpyAsync(profiles_out, d_prof, (size_t)kProfiles * nq * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PRIPPS_OK;
}

// ---------------------------------------------------------------------------
// model queries
// ---------------------------------------------------------------------------
extern "C" int pripps_debye_model_nq(const pripps_debye_model *m, int32_t *nq) {
    if (!m || !nq) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    *nq = m->nq;
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_q_values(const pripps_debye_model *m, double *out) {
    if (!m || (!out && m->nq)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    std::copy(m->q.begin(), m->q.end(), out);
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_intensity(const pripps_debye_model *m, double vs, double cd,
                                            double *total, double *atoms, double *excluded,
                                            double *hydration) {
    if (!m) return fail(PRIPPS_ERR_USAGE, "model is NULL");
    if (m->nq == 0) return PRIPPS_OK;
    pripps_engine *e = m->engine;
    std::vector<double> host((size_t)4 * m->nq);
    {
        std::lock_guard<std::mutex> lock(e->mu);
        HIPCHK(hipSetDevice(e->device));
        HIPCHK(launch_combine(m->d_prof(), m->nq, vs, cd, m->d_out(), e->stream));
        HIPCHK(hipMemcpyAsync(host.data(), m->d_out(), host.size() * 8, hipMemcpyDeviceToHost,
                              e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    double *dst[4] = {total, atoms, excluded, hydration};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) std::copy(host.begin() + (size_t)k * m->nq, host.begin() + (size_t)(k + 1) * m->nq, dst[k]);
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_profiles(const pripps_debye_model *m, double *out) {
    if (!m || (!out && m->nq)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (m->nq == 0) return PRIPPS_OK;
    pripps_engine *e = m->engine;
    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(out, m->d_prof(), (size_t)6 * m->nq * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_fit(const pripps_debye_model *m, const double *q,
                                      const double *inten, const double *sigma, int64_t n,
                                      int32_t method, int32_t c1_points, int32_t c2_points,
                                      const double *c1_bounds, const double *c2_bounds,
                                      pripps_fit_result *out) {
    if (!m || !out) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (n <= 0) return fail(PRIPPS_ERR_VALIDATION, "experimental data is empty");
    if (!q || !inten) return fail(PRIPPS_ERR_USAGE, "q and intensity must be non-NULL");
    if (method != PRIPPS_FIT_NELDER_MEAD && method != PRIPPS_FIT_GRID)
        return fail(PRIPPS_ERR_USAGE, "unknown fit method " + std::to_string(method));
    const double c1d[2] = {0.5, 2.0}, c2d[2] = {0.0, 0.1};
    const double *c1b = c1_bounds ? c1_bounds : c1d;
    const double *c2b = c2_bounds ? c2_bounds : c2d;
    if (method == PRIPPS_FIT_GRID && (c1_points < 2 || c2_points < 2))
        return fail(PRIPPS_ERR_VALIDATION,
                    "grid fit requires at least 2 points per axis (got c1_points=" +
                        std::to_string(c1_points) + ", c2_points=" + std::to_string(c2_points) + ")");
    // the device fit indexes grid points and data points with 32-bit ints
    if (method == PRIPPS_FIT_GRID && (int64_t)c1_points * (int64_t)c2_points > INT32_MAX)
        return fail(PRIPPS_ERR_VALIDATION, "grid fit: c1_points * c2_points exceeds 2^31 - 1");
    if (n > INT32_MAX) return fail(PRIPPS_ERR_VALIDATION, "experimental data exceeds 2^31 - 1 points");
    const double *bb[2] = {c1b, c2b};
    const char *nm[2] = {"c1", "c2"};
    for (int k = 0; k < 2; ++k) {
        if (!std::isfinite(bb[k][0]) || !std::isfinite(bb[k][1]))
            return fail(PRIPPS_ERR_VALIDATION, std::string(nm[k]) + " bounds must be finite");
        if (bb[k][0] >= bb[k][1])
            return fail(PRIPPS_ERR_VALIDATION,
                        std::string(nm[k]) + " lower bound must be less than upper bound");
    }
    if (m->nq == 0) return fail(PRIPPS_ERR_VALIDATION, "model has no q points");
    pripps_engine *e = m->engine;
    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));
    int rc = ensure_fit(e, device_fit_doubles(m->nq, n, method, c1_points, c2_points));
    if (rc) return rc;
    std::string err;
    rc = device_fit(e->stream, e->fit_buf, m->d_prof(), m->d_q(), m->nq, q, inten, sigma, n,
                    method, c1_points, c2_points, c1b, c2b, out, &err);
    trim_fit(e);
    if (rc) return fail(rc, err);
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_from_profiles(pripps_engine *e, const double *q, int32_t nq,
                                                const double *profiles,
                                                pripps_debye_model **out) {
    if (!e || !out) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    *out = nullptr;
    if (nq < 0) return fail(PRIPPS_ERR_VALIDATION, "nq must be >= 0");
    if (nq > 0 && (!q || !profiles)) return fail(PRIPPS_ERR_USAGE, "q and profiles must be non-NULL");
    if (!finite_all(q, nq) || !finite_all(profiles, (size_t)6 * nq))
        return fail(PRIPPS_ERR_VALIDATION, "q and profiles must be finite");
    auto *m = new pripps_debye_model();
    m->engine = e;
    m->nq = nq;
    m->q.assign(q, q + nq);
    {
        std::lock_guard<std::mutex> lock(e->mu);
        hipError_t err = hipSetDevice(e->device);
        if (err == hipSuccess) err = model_alloc(m);
        if (err == hipSuccess && nq > 0)
            err = hipMemcpyAsync(m->d_prof(), profiles, (size_t)6 * nq * 8, hipMemcpyHostToDevice, e->stream);
        if (err == hipSuccess && nq > 0)
            err = hipMemcpyAsync(m->d_q(), q, (size_t)nq * 8, hipMemcpyHostToDevice, e->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
        if (err != hipSuccess) {
            (void)hipStreamSynchronize(e->stream);
            model_release_block(m);
            delete m;
            return hip_fail(err, "pripps_debye_model_from_profiles");
        }
    }
    e->refs.fetch_add(1);
    *out = m;
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_stats(const pripps_debye_model *m, pripps_build_stats *out) {
    if (!m || !out) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    *out = m->stats;
    return PRIPPS_OK;
}

// The block returns to the pool without a device-wide wait.  Work that reads
// it is on the engine's stream; a later model that receives the block writes
// it only from a later submission on a stream whose work is ordered after
// this call by the caller (every query of a model synchronises its stream
// before returning), so the block is idle when it is handed out again.
extern "C" int pripps_debye_model_free(pripps_debye_model *m) {
    if (!m) return PRIPPS_OK;
    pripps_engine *e = m->engine;
    model_release_block(m);
    delete m;
    engine_release(e);
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_fit_curve(const pripps_debye_model *m, double vs, double cd,
                                            const double *q, int64_t n, double *out) {
    if (!m) return fail(PRIPPS_ERR_USAGE, "model is NULL");
    if (n < 0) return fail(PRIPPS_ERR_VALIDATION, "n must be >= 0");
    if (n > 0 && (!q || !out)) return fail(PRIPPS_ERR_USAGE, "q and out must be non-NULL");
    if (n == 0) return PRIPPS_OK;
    if (m->nq == 0) return fail(PRIPPS_ERR_VALIDATION, "model has no q points");
    if (!finite_all(q, (size_t)n)) return fail(PRIPPS_ERR_VALIDATION, "q must be finite");
    pripps_engine *e = m->engine;
    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));
    int rc = ensure_fit(e, device_spline_total_doubles(m->nq, n));
    if (rc) return rc;
    std::string err;
    rc = device_spline_total(e->stream, e->fit_buf, m->d_prof(), m->d_q(), m->nq, vs, cd, q, n,
                             out, &err);
    trim_fit(e);
    if (rc) return fail(rc, err);
    return PRIPPS_OK;
}

// The engine's most recent launches as one instantiated HIP graph, captured at
// the first graph replay of a build and kept until the next build.  The
// launches hold kernels only (no memset nodes): the Chebyshev kernel leaves
// its queue counters at zero and the recurrence kernels' counter is zeroed
// by a fill kernel.  The caller holds e->mu with e's device current.
static int ensure_replay_graph(pripps_engine *e) {
    if (e->replay_exec && e->replay_serial == e->build_serial) return PRIPPS_OK;
    if (e->replay_exec) (void)hipGraphExecDestroy(e->replay_exec);
    if (e->replay_graph) (void)hipGraphDestroy(e->replay_graph);
    e->replay_exec = nullptr;
    e->replay_graph = nullptr;
    hipStream_t st = e->stream;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (const auto &r : e->last_launches)
        if (err == hipSuccess) err = run_launch(r, st, nullptr, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t eerr = hipStreamEndCapture(st, &g);
    if (err == hipSuccess) err = eerr;
    if (err == hipSuccess) err = hipGraphInstantiate(&e->replay_exec, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        e->replay_exec = nullptr;
        return hip_fail(err, "capture of the replay graph");
    }
    e->replay_graph = g;
    e->replay_serial = e->build_serial;
    return PRIPPS_OK;
}

// Replay the engine's most recent launches `repeats` times; the caller holds
// e->mu.  With pair_ms_mean: eager launches with HIP events around each pair
// kernel, pair_ms_mean = their mean device time per replay.  Without: one
// launch of the captured graph per replay.  total_ms: device time of all
// replays.
static int replay_timed(pripps_engine *e, int32_t repeats, double *pair_ms_mean, double *total_ms) {
    HIPCHK(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    if (!pair_ms_mean) {
        int rc = ensure_replay_graph(e);
        if (rc) return rc;
        hipError_t err = hipEventRecord(e->ev0, st);
        for (int r = 0; r < repeats && err == hipSuccess; ++r) err = hipGraphLaunch(e->replay_exec, st);
        if (err == hipSuccess) err = hipEventRecord(e->ev1, st);
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        if (err != hipSuccess) return hip_fail(err, "graph replay");
        float tot = 0;
        (void)hipEventElapsedTime(&tot, e->ev0, e->ev1);
        if (total_ms) *total_ms = tot;
        return PRIPPS_OK;
    }
    const size_t nl = e->last_launches.size();
    std::vector<hipEvent_t> ev(2 * (size_t)repeats * nl + 2, nullptr);
    for (auto &x : ev) HIPCHK(hipEventCreate(&x));
    auto destroy = [&]() {
        for (auto &x : ev)
            if (x) (void)hipEventDestroy(x);
    };
    hipError_t err = hipEventRecord(ev[0], st);
    size_t k = 2;
    for (int r = 0; r < repeats && err == hipSuccess; ++r)
        for (size_t l = 0; l < nl && err == hipSuccess; ++l, k += 2)
            err = run_launch(e->last_launches[l], st, ev[k], ev[k + 1]);
    if (err == hipSuccess) err = hipEventRecord(ev[1], st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) {
        destroy();
        return hip_fail(err, "replay");
    }
    double sum = 0;
    for (size_t x = 2; x < k; x += 2) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[x], ev[x + 1]);
        sum += ms;
    }
    float tot = 0;
    (void)hipEventElapsedTime(&tot, ev[0], ev[1]);
    destroy();
    if (pair_ms_mean) *pair_ms_mean = repeats ? sum / repeats : 0.0;
    if (total_ms) *total_ms = tot;
    return PRIPPS_OK;
}

extern "C" int pripps_debye_model_recompute(pripps_debye_model *m, int32_t repeats,
                                            double *pair_ms_mean, double *total_ms) {
    if (!m) return fail(PRIPPS_ERR_USAGE, "model is NULL");
    if (repeats < 0) return fail(PRIPPS_ERR_USAGE, "repeats must be >= 0");
    pripps_engine *e = m->engine;
    std::lock_guard<std::mutex> lock(e->mu);
    if (m->build_serial == 0 || m->build_serial != e->build_serial || e->last_plain)
        return fail(PRIPPS_ERR_USAGE,
                    "recompute needs a tiled model from the engine's most recent build");
    return replay_timed(e, repeats, pair_ms_mean, total_ms);
}

// ---------------------------------------------------------------------------
// autotuning
// ---------------------------------------------------------------------------
static void put_entry(pripps_engine *e, const pripps_tuning_entry &x) {
    std::lock_guard<std::mutex> lk(e->tune_mu);
    for (auto &y : e->tune_table)
        if (y.precision == x.precision && y.uniform == x.uniform && y.log2_atoms == x.log2_atoms &&
            y.log2_nq == x.log2_nq) {
            y = x;
            return;
        }
    e->tune_table.push_back(x);
}

extern "C" int pripps_engine_autotune(pripps_engine *e, const pripps_debye_input *in,
                                      int32_t precision, int32_t repeats, pripps_tuning *best,
                                      double *best_pair_ms) {
    if (!e) return fail(PRIPPS_ERR_USAGE, "engine is NULL");
    if (!in) return fail(PRIPPS_ERR_USAGE, "input is NULL");
    if (precision != PRIPPS_FP32_COMPENSATED && precision != PRIPPS_FP64)
        return fail(PRIPPS_ERR_USAGE, "autotune: precision must be FP32_COMPENSATED or FP64");
    repeats = std::max(1, std::min(repeats, 100));
    Prepared p;
    int rc = prepare(in, p);
    if (rc) return rc;
    if (p.n < 2 || p.nq < 1) return fail(PRIPPS_ERR_VALIDATION, "autotune needs >= 2 atoms and >= 1 q");

    struct Cand {
        int Q, shift, block, variant;
        bool operator==(const Cand &o) const {
            return Q == o.Q && shift == o.shift && block == o.block && variant == o.variant;
        }
    };
    std::vector<std::pair<Cand, double>> seen;
    // block 0: the plan's default for Q; a Chebyshev plan at shift 0 sizes
    // its units itself
    auto absolute = [&](const Cand &c) {
        pripps_tuning t{};
        t.q_per_lane = c.Q;
        t.block = c.block > 0 ? c.block : c.Q == kChebQ ? kChebBlock : kDefaultBlock;
        t.span = c.Q == kChebQ && c.shift == 0 ? 0 : shifted_span(default_span(p, c.Q, e->n_cu), c.shift);
        t.variant = c.variant;
        return t;
    };
    auto eval = [&](const Cand &c, double &ms) -> int {
        for (auto &x : seen)
            if (x.first == c) {
                ms = x.second;
                return PRIPPS_OK;
            }
        const pripps_tuning t = absolute(c);
        pripps_debye_model *m = nullptr;
        int r = build_models(e, 1, in, precision, &t, &m);
        if (r) return r;
        {
            std::lock_guard<std::mutex> lock(e->mu);
            r = replay_timed(e, repeats, &ms, nullptr);
        }
        pripps_debye_model_free(m);
        if (r) return r;
        seen.push_back({c, ms});
        return PRIPPS_OK;
    };

    Cand cur{default_q(precision, p.uniform, p.nq, p.n), 0, 0, 3};
    double cur_ms = 0;
    rc = eval(cur, cur_ms);
    if (rc) return rc;
    auto try_all = [&](std::vector<Cand> cands) -> int {
        for (const Cand &c : cands) {
            double ms = 0;
            int r = eval(c, ms);
            if (r) return r;
            if (ms < cur_ms) {
                cur = c;
                cur_ms = ms;
            }
        }
        return PRIPPS_OK;
    };
    // 1. q points per lane
    {
        std::vector<Cand> c;
        const int qs[4] = {32, 64, 128, kChebQ};
        for (int Q : qs)
            if (valid_q(precision, p.uniform, Q) == Q) c.push_back({Q, 0, cur.block, 3});
        if ((rc = try_all(c))) return rc;
    }
    // 2. j-span
    {
        std::vector<Cand> c;
        for (int sh : {-2, -1, 1, 2}) c.push_back({cur.Q, sh, cur.block, cur.variant});
        if ((rc = try_all(c))) return rc;
    }
    // 3. pair block
    {
        std::vector<Cand> c;
        // capped at 128: longer blocks trade accuracy on lattice-placed points
        for (int b : {16, 32, 64, 128}) c.push_back({cur.Q, cur.shift, b, cur.variant});
        if ((rc = try_all(c))) return rc;
    }
    // 4. inner-loop variant (FP32, uniform grids)
    {
        std::vector<Cand> c;
        for (int v : {1, 4})
            if (valid_variant(precision, p.uniform, cur.Q, v) == v) c.push_back({cur.Q, cur.shift, cur.block, v});
        if ((rc = try_all(c))) return rc;
    }
    pripps_tuning_entry x{};
    x.precision = precision;
    x.uniform = p.uniform ? 1 : 0;
    x.log2_atoms = log2_atoms_of(p.n);
    x.log2_nq = log2_nq_of(p.nq);
    x.q_per_lane = cur.Q;
    x.block = cur.block;
    x.span_shift = cur.shift;
    x.variant = cur.variant;
    x.pair_ms = cur_ms;
    put_entry(e, x);
    if (best) *best = absolute(cur);
    if (best_pair_ms) *best_pair_ms = cur_ms;
    return PRIPPS_OK;
}

extern "C" int pripps_engine_tuning_table(pripps_engine *e, pripps_tuning_entry *entries,
                                          int32_t capacity, int32_t *count) {
    if (!e || !count) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (capacity > 0 && !entries) return fail(PRIPPS_ERR_USAGE, "entries is NULL");
    std::lock_guard<std::mutex> lk(e->tune_mu);
    const int32_t n = (int32_t)e->tune_table.size();
    for (int32_t k = 0; k < std::min(n, capacity); ++k) entries[k] = e->tune_table[k];
    *count = n;
    return PRIPPS_OK;
}

extern "C" int pripps_engine_set_tuning_entry(pripps_engine *e, const pripps_tuning_entry *x) {
    if (!e || !x) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (x->precision != PRIPPS_FP32_COMPENSATED && x->precision != PRIPPS_FP64)
        return fail(PRIPPS_ERR_VALIDATION, "tuning entry: precision must be FP32_COMPENSATED or FP64");
    if (x->log2_atoms < 0 || x->log2_atoms > 31 || x->log2_nq < 0 || x->log2_nq > 30 ||
        x->span_shift < -8 || x->span_shift > 8 || x->block < 0 || x->block > 4096)
        return fail(PRIPPS_ERR_VALIDATION, "tuning entry out of range");
    if (valid_q(x->precision, x->uniform != 0, x->q_per_lane) != x->q_per_lane)
        return fail(PRIPPS_ERR_VALIDATION,
                    "tuning entry: q_per_lane " + std::to_string(x->q_per_lane) +
                        " is not available for this precision and grid");
    pripps_tuning_entry y = *x;
    y.uniform = x->uniform ? 1 : 0;
    put_entry(e, y);
    return PRIPPS_OK;
}

// ---------------------------------------------------------------------------
// trajectory frames
// ---------------------------------------------------------------------------
extern "C" int pripps_debye_frames_average(pripps_engine *e, const pripps_debye_input *topo,
                                           int32_t n_frames, const double *coords,
                                           const double *w, int32_t stride,
                                           int32_t precision, const pripps_tuning *tune,
                                           int32_t max_frames_per_launch,
                                           pripps_debye_model **out) {
    const auto t_start = std::chrono::steady_clock::now();
    if (!out) return fail(PRIPPS_ERR_USAGE, "out is NULL");
    *out = nullptr;
    if (!e || !topo) return fail(PRIPPS_ERR_USAGE, "engine and topology must be non-NULL");
    if (precision != PRIPPS_FP32_COMPENSATED && precision != PRIPPS_FP64)
        return fail(PRIPPS_ERR_USAGE, "frames: precision must be FP32_COMPENSATED or FP64");
    if (n_frames < 0) return fail(PRIPPS_ERR_VALIDATION, "n_frames must be >= 0");
    if (stride < 1) return fail(PRIPPS_ERR_VALIDATION, "stride must be >= 1");
    if (max_frames_per_launch < 0) return fail(PRIPPS_ERR_VALIDATION, "max_frames_per_launch must be >= 0");
    const int64_t n = topo->n_atoms;
    if (n_frames > 0 && n > 0 && !coords) return fail(PRIPPS_ERR_USAGE, "coords is NULL");
    // weights: the file's total must be usable (read_weights), and so must
    // the total of the frames the stride keeps (trajectory_average)
    if (w) {
        double sum = 0;
        for (int32_t f = 0; f < n_frames; ++f) sum += w[f];
        if (!std::isnormal(std::fabs(sum)))
            return fail(PRIPPS_ERR_VALIDATION, "sum of weights is zero or infinite");
    }
    std::vector<int32_t> used;
    double wsum = 0;
    for (int32_t f = 0; f < n_frames; f += stride) {
        used.push_back(f);
        wsum += w ? w[f] : 1.0;
    }
    if (!std::isnormal(std::fabs(wsum))) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", wsum);
        return fail(PRIPPS_ERR_VALIDATION, "weights of the " + std::to_string(used.size()) +
                                               " averaged frames sum to " + buf +
                                               "; cannot normalize");
    }
    const int32_t nq = topo->nq;
    const int M = (int)used.size();

    // per-frame host preparation (SoA coordinates of frame f)
    std::vector<double> fx((size_t)n), fy((size_t)n), fz((size_t)n);
    auto prep_frame = [&](int32_t f, Prepared &P) -> int {
        const double *c = coords + (size_t)f * (size_t)n * 3;
        for (int64_t i = 0; i < n; ++i) {
            fx[i] = c[3 * i];
            fy[i] = c[3 * i + 1];
            fz[i] = c[3 * i + 2];
        }
        pripps_debye_input in = *topo;
        in.x = fx.data();
        in.y = fy.data();
        in.z = fz.data();
        int rc = prepare(&in, P);
        if (rc) g_last_error = "frame " + std::to_string(f) + ": " + g_last_error;
        return rc;
    };

    // Every chunk shares the first full chunk's plan (and the captured graph
    // relies on that), and a Chebyshev plan's strips depend on the largest
    // pair distance: plan every frame with one bound that holds for all used
    // frames (twice the largest distance from a frame's centroid).
    double rmax_all = 0;
    for (int32_t f : used) {
        const double *c = coords + (size_t)f * (size_t)n * 3;
        double cx = 0, cy = 0, cz = 0;
        for (int64_t i = 0; i < n; ++i) {
            cx += c[3 * i];
            cy += c[3 * i + 1];
            cz += c[3 * i + 2];
        }
        if (n > 0) {
            cx /= (double)n;
            cy /= (double)n;
            cz /= (double)n;
        }
        double r2 = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double dx = c[3 * i] - cx, dy = c[3 * i + 1] - cy, dz = c[3 * i + 2] - cz;
            r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
        }
        rmax_all = std::max(rmax_all, 2.0 * std::sqrt(r2) * (1.0 + 1e-6) + 1e-6);
    }
    // chunk size: frames per persistent launch, bounded by a 4 GiB workspace
    Prepared p0;
    int rc = prep_frame(used[0], p0);
    if (rc) return rc;
    p0.rmax = std::max(p0.rmax, rmax_all);
    if (nq == 0 || p0.n == 0) return fail(PRIPPS_ERR_VALIDATION, "frames need atoms and q points");
    int C = M;
    {
        BuildPlan B1;
        std::vector<Prepared> one(1, p0);
        plan_build(e, one, precision, tune, (size_t)nq * 8 + 8, B1);
        const size_t per_frame = B1.total_bytes + (size_t)kProfiles * nq * 8;
        const size_t cap = (size_t)4 << 30;
        C = (int)std::max<size_t>(1, std::min<size_t>((size_t)M, cap / std::max<size_t>(1, per_frame)));
    }
    // at least four chunks (of >= 16 frames): the host side of every chunk
    // but the first then overlaps the kernels of the one before
    C = std::min(C, std::max(16, (M + 3) / 4));
    if (max_frames_per_launch > 0) C = std::min(C, (int)max_frames_per_launch);

    std::lock_guard<std::mutex> lock(e->mu);
    HIPCHK(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    // the arena is overwritten: nothing of an earlier build stays replayable
    e->build_serial++;
    e->last_launches.clear();

    auto *model = new pripps_debye_model();
    model->engine = e;
    model->nq = nq;
    model->q = p0.q;
    double *work = nullptr;  // [C][6 nq] frame profiles | [6 nq] accumulator
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);  // nothing in flight may use what is freed
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (work) (void)hipFree(work);
        if (model) {
            model_release_block(model);
            delete model;
        }
    };
#define FRAMES_CHK(expr)                     \
    do {                                     \
        hipError_t e_ = (expr);              \
        if (e_ != hipSuccess) {              \
            cleanup();                       \
            return hip_fail(e_, #expr);      \
        }                                    \
    } while (0)
    const size_t prof_n = (size_t)kProfiles * nq;
    FRAMES_CHK(model_alloc(model));
    FRAMES_CHK(hipMalloc((void **)&work, ((size_t)C + 1) * prof_n * sizeof(double)));
    double *acc = work + (size_t)C * prof_n;
    FRAMES_CHK(launch_fill_double(acc, (int64_t)prof_n, 0.0, st));

    double kernel_ms = 0;
    int64_t items_per_chunk = 0;
    int graph_launches = 0;
    StructPlan sp0;
    // Host side of one chunk: its frames prepared (in parallel over frames;
    // prepare() is a pure function of its input) and planned.  The next
    // chunk's host side runs while the current chunk's kernels do; only the
    // pinned staging buffer has to wait for the previous upload.  A plan
    // depends on the topology, the q grid and the frame count only, so every
    // full chunk shares the first one's (the captured graph relies on that
    // too); only the remainder chunk is planned anew.
    struct ChunkHost {
        int first = 0, cnt = 0;
        std::vector<Prepared> P;
        std::shared_ptr<const BuildPlan> B;
    };
    std::shared_ptr<const BuildPlan> full_plan;
    const int host_threads =
        std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    auto prep_chunk = [&](int first, ChunkHost &H) -> int {
        H.first = first;
        H.cnt = std::min(C, M - first);
        H.P.assign(H.cnt, Prepared{});
        H.B.reset();
        const int nt = std::max(1, std::min(H.cnt, host_threads));
        std::vector<int> rcs(H.cnt, PRIPPS_OK);
        std::vector<std::string> msgs(H.cnt);
        auto part = [&](int t) {
            std::vector<double> bx((size_t)n), by((size_t)n), bz((size_t)n);
            for (int c = t; c < H.cnt; c += nt) {
                if (first + c == 0) {
                    H.P[c] = p0;
                    continue;
                }
                const int32_t f = used[first + c];
                const double *cc = coords + (size_t)f * (size_t)n * 3;
                for (int64_t i = 0; i < n; ++i) {
                    bx[i] = cc[3 * i];
                    by[i] = cc[3 * i + 1];
                    bz[i] = cc[3 * i + 2];
                }
                pripps_debye_input in = *topo;
                in.x = bx.data();
                in.y = by.data();
                in.z = bz.data();
                rcs[c] = prepare(&in, H.P[c]);
                H.P[c].rmax = std::max(H.P[c].rmax, rmax_all);
                if (rcs[c]) {
                    msgs[c] = "frame " + std::to_string(f) + ": " + g_last_error;  // this thread's
                } else if (H.P[c].n != p0.n || H.P[c].nq != p0.nq) {
                    rcs[c] = PRIPPS_ERR_VALIDATION;
                    msgs[c] = "frame topology changed";
                }
            }
        };
        if (nt == 1) {
            part(0);
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < nt; ++t) pool.emplace_back(part, t);
            for (auto &th : pool) th.join();
        }
        for (int c = 0; c < H.cnt; ++c)
            if (rcs[c]) {
                g_last_error = msgs[c];
                return rcs[c];
            }
        if (H.cnt == C && full_plan) {
            H.B = full_plan;
        } else {
            auto b = std::make_shared<BuildPlan>();
            plan_build(e, H.P, precision, tune, (size_t)nq * 8 + (size_t)H.cnt * 8, *b);
            H.B = b;
            if (H.cnt == C) full_plan = H.B;
        }
        return PRIPPS_OK;
    };
    ChunkHost cur, next;
    rc = prep_chunk(0, cur);
    if (rc) {
        cleanup();
        return rc;
    }
    // size the arena and staging once for the largest (first) chunk, so the
    // captured graph's pointers stay valid for every replay
    rc = ensure_arena(e, cur.B->total_bytes);
    if (!rc) rc = ensure_staging(e, cur.B->upload_bytes);
    if (rc) {
        cleanup();
        return rc;
    }
    sp0 = cur.B->sp[0];
    items_per_chunk = cur.B->n_items;
    const BuildPlan *staged = nullptr;  // the plan whose work lists the staging buffer holds
    for (;;) {
        const BuildPlan &B = *cur.B;
        const std::vector<Prepared> &P = cur.P;
        const int cnt = cur.cnt, first = cur.first;
        char *dbase = (char *)e->arena;
        char *hbase = (char *)e->staging;
        const double *d_q = (const double *)(dbase + B.o_user);
        const double *d_w = d_q + nq;
        std::vector<const double *> dq(cnt, d_q);
        std::vector<double *> dprof(cnt);
        for (int c = 0; c < cnt; ++c) dprof[c] = work + (size_t)c * prof_n;
        stage_build(B, P, hbase, dbase, dq, dprof, staged != &B);
        staged = &B;
        std::memcpy(hbase + B.o_user, p0.q.data(), (size_t)nq * 8);
        double *hw = (double *)(hbase + B.o_user) + nq;
        for (int c = 0; c < cnt; ++c) hw[c] = w ? w[used[first + c]] : 1.0;
        std::vector<LaunchRecord> launches;
        rc = make_launches(e, B, dbase, launches);
        if (rc) {
            cleanup();
            return rc;
        }
        auto enqueue_all = [&]() -> hipError_t {
            hipError_t err = hipMemcpyAsync(dbase, hbase, B.upload_bytes, hipMemcpyHostToDevice, st);
            if (err == hipSuccess) err = enqueue_kernels(B, P, dbase, launches, st);
            if (err == hipSuccess) err = launch_frames_accumulate(work, cnt, d_w, nq, acc, st);
            return err;
        };
        const bool full = cnt == C;
        if (full && !gexec) {
            // capture the submission of a full chunk once; later full chunks
            // replay it with their own frames in the same staging layout
            FRAMES_CHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipError_t cerr = enqueue_all();
            hipError_t eerr = hipStreamEndCapture(st, &graph);
            if (cerr != hipSuccess || eerr != hipSuccess) {
                cleanup();
                return hip_fail(cerr != hipSuccess ? cerr : eerr, "frames: graph capture");
            }
            FRAMES_CHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        FRAMES_CHK(hipEventRecord(e->ev0, st));
        if (full) {
            FRAMES_CHK(hipGraphLaunch(gexec, st));
            ++graph_launches;
        } else {
            FRAMES_CHK(enqueue_all());
        }
        FRAMES_CHK(hipEventRecord(e->ev1, st));
        // the next chunk's host side overlaps this chunk's kernels
        const int nfirst = first + cnt;
        if (nfirst < M) {
            rc = prep_chunk(nfirst, next);
            if (rc) {
                cleanup();
                return rc;
            }
        }
        FRAMES_CHK(hipStreamSynchronize(st));  // staging is rewritten next chunk
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
        kernel_ms += ms;
        if (nfirst >= M) break;
        std::swap(cur, next);
    }
    FRAMES_CHK(launch_frames_normalize(acc, nq, wsum, model->d_prof(), st));
    FRAMES_CHK(hipMemcpyAsync(model->d_q(), p0.q.data(), (size_t)nq * 8, hipMemcpyHostToDevice, st));
    FRAMES_CHK(hipStreamSynchronize(st));
#undef FRAMES_CHK
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipFree(work);
    e->refs.fetch_add(1);
    pripps_build_stats &S = model->stats;
    S.kernel_ms = kernel_ms;
    S.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    S.pair_q_terms = p0.pair_q_terms * (double)M;
    S.q_per_lane = sp0.Q;
    S.block = sp0.block;
    S.span = sp0.span;
    S.work_items = items_per_chunk;
    S.frames = M;
    S.frames_per_launch = C;
    S.graph_launches = graph_launches;
    S.cheb_order = sp0.P;
    *out = model;
    return PRIPPS_OK;
}

namespace pripps {
// used by the serving loop
int engine_build(pripps_engine *e, int32_t n, const pripps_debye_input *inputs, int32_t precision,
                 const pripps_tuning *tuning, pripps_debye_model **out) {
    return build_models(e, n, inputs, precision, tuning, out);
}
void set_last_error(const std::string &msg) { g_last_error = msg; }

// Device bytes one structure needs when built alone on `e`: the arena of its
// plan (atoms, tables, work lists, item partials, chunk totals) and its model
// block.  The plan is the one the build would use (same tuning table), so
// the figure is exact for a build alone and a close upper bound per member
// of a batch (only alignment padding is shared).  Returns false if the input
// does not prepare (the build will report why).
bool engine_estimate(pripps_engine *e, const pripps_debye_input *in, int32_t precision,
                     uint64_t *arena_bytes, uint64_t *model_block) {
    std::vector<Prepared> P(1);
    const std::string saved = g_last_error;
    const int rc = prepare(in, P[0]);
    g_last_error = saved;
    if (rc) return false;
    BuildPlan B;
    plan_build(e, P, precision, nullptr, 0, B);
    *arena_bytes = B.total_bytes;
    *model_block = (uint64_t)1 << pool_class(model_bytes(P[0].nq));
    return true;
}

// Device bytes of e's arena (grow-only).
uint64_t engine_arena_bytes(pripps_engine *e) { return e->arena_pub.load(); }

// Free e's arena; the next build allocates a new one.  Called only while no
// build runs on e.
int engine_trim_arena(pripps_engine *e) {
    std::lock_guard<std::mutex> lock(e->mu);
    if (!e->arena) return PRIPPS_OK;
    HIPCHK(hipSetDevice(e->device));
    hipError_t err = hipFree(e->arena);
    e->arena = nullptr;
    e->arena_size = 0;
    e->arena_pub = 0;
    e->build_serial++;  // nothing of the freed arena is replayable
    e->last_launches.clear();
    if (err != hipSuccess) return hip_fail(err, "hipFree(arena)");
    return PRIPPS_OK;
}
const std::string &get_last_error() { return g_last_error; }
// used by the serving loop: a model built on a lane is handed to another
// engine of the same device, so that its queries (intensity, profiles, fit)
// run on that engine's stream under its lock instead of queueing behind the
// lane's next batch.  The build has completed (build_models synchronises the
// lane's stream), so the profiles are visible to every stream of the device.
// The model's kernels are no longer the new engine's last build: recompute
// refuses it.
void model_rehome(pripps_debye_model *m, pripps_engine *to) {
    if (!m || !to || m->engine == to) return;
    to->refs.fetch_add(1);
    pripps_engine *from = m->engine;
    m->engine = to;
    m->build_serial = 0;
    engine_release(from);
}

// used by the Direct transform (direct.hip): the engine's device, stream and
// submission lock; a model keeps its engine alive through retain/release.
EngineAccess engine_access(pripps_engine *e) {
    return EngineAccess{e->device, e->n_cu, e->stream, e->ev0, e->ev1, &e->mu};
}
void engine_retain(pripps_engine *e) { e->refs.fetch_add(1); }
void engine_unref(pripps_engine *e) { engine_release(e); }
}  // namespace pripps
