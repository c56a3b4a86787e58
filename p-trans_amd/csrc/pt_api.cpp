// pt_api.cpp -- the C-ABI of libptrans (include/ptrans.h): contexts, geometry and
// mode tables in HBM, and the trace entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ptrans.h"
#include "pt_internal.h"
#include "pt_philox.h"
#include "pt_math.h"

using namespace pt;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define PT_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(PT_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                        __FILE__, __LINE__);                                            \
    } while (0)

template <class T>
int upload(const std::vector<T> &v, T **dptr, hipStream_t s)
{
    *dptr = nullptr;
    if (v.empty()) return PT_OK;
    PT_HIP(hipMalloc((void **)dptr, v.size() * sizeof(T)));
    PT_HIP(hipMemcpyAsync(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return PT_OK;
}

constexpr double kPi = 3.14159265358979323846;

// Scratch of the allocation-free alias trace (pt_trace_alias_noalloc): the queue and
// statistics counters, the per-mode tallies and the per-branch scales of one call.  Each
// device holds its own copy of the array from the library's load until the process ends;
// its device address is looked up once per device when a context is created there.
constexpr size_t kNoAllocMaxModes = 65536;
constexpr int kNoAllocMaxBranches = 64;
constexpr size_t kNoAllocCounters = kCounterWords;
// then the records of the paths handed to the tail kernel (16-byte aligned)
constexpr size_t kNoAllocTailWord = (kNoAllocCounters + 2 * kNoAllocMaxModes + kNoAllocMaxBranches + 1) & ~(size_t)1;
constexpr int64_t kNoAllocTailRecs = 8192;
constexpr size_t kNoAllocWords = kNoAllocTailWord + kNoAllocTailRecs * sizeof(TailRec) / 8;
static_assert(kNoAllocMaxModes == PT_NOALLOC_MAX_MODES && kNoAllocMaxBranches == PT_NOALLOC_MAX_BRANCHES,
              "limits of the allocation-free trace");
static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_addr[kMaxDevices] = {};   // device address of `scratch` per device

// The device address of `scratch` on the current device `device` (looked up once).
hipError_t scratch_address(int device, unsigned long long **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = (unsigned long long *)p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

}  // namespace

struct pt_ctx {
    int device = 0;
    uint64_t seed = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    bool count_ops = false;
    int ncu = 0;
    int blocks_per_cu = 0;                     // 0: occupancy maximum of each kernel
    int threads = 256;
    int chunk = 256;
    int refill_min = 16;
    int64_t event_limit = 100000000;
    // grow-only scratch of the grid trace: outcome bytes (padded by 16 for the reduction's
    // 16-byte loads), and the angle tables and cell counts of a call
    uint8_t *outcome = nullptr;
    size_t outcome_bytes = 0;
    void *grid_buf = nullptr;
    size_t grid_buf_bytes = 0;
    void *host_stage = nullptr;                // pinned host copy of the cell counts
    size_t host_stage_bytes = 0;
    unsigned long long *counters = nullptr;   // queue + 6 stats + skipped
    double *scale = nullptr;                   // per-branch scale of the current call
    int scale_n = 0;
    unsigned long long *noalloc = nullptr;     // this device's `scratch` (device address)
    // grow-only work space of the alias, adjoint and grain-tally calls (tallies, groups,
    // sums), kept between calls so a repeated call allocates nothing
    void *work = nullptr;
    size_t work_bytes = 0;
    // grow-only buffer of the paths handed to the tail kernel, and their adjoint scores
    TailRec *tail = nullptr;
    int64_t tail_alloc = 0;
    double *tail_score = nullptr;
    int64_t tail_score_alloc = 0;
    int64_t tail_events = 500;                 // a path past this many events is handed over
    int64_t tail_max = 1 << 20;                // at most this many records per launch
    uint32_t tail_epoch = 0;                   // launches that wrote tail records so far
    // the tail kernel beside the persistent one: its workgroups (0: none), stream, events
    int live_blocks = 16;
    // grow-only list of the paths the queue runs first (alias-mode launches with a tail)
    int64_t *first_ids = nullptr;
    int64_t first_alloc = 0;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int64_t live_ticks = 0;                    // 20 s of the device's wall clock
    size_t lds_cu = 0;                         // LDS bytes per CU
};

struct pt_geom {
    pt_ctx *ctx = nullptr;
    DevGeom dg{};
    std::vector<DSurf> surf;        // host copy of the surface records
    std::vector<void *> owned;
};

struct pt_modes {
    pt_ctx *ctx = nullptr;
    DevModes dm{};
    int32_t nb = 0, nk = 0;
    bool has_alias = false;
    std::vector<void *> owned;
};

extern "C" {

int32_t pt_version(void) { return PT_ABI_VERSION; }

int pt_philox_fill(uint64_t seed, uint64_t path, uint64_t first_draw, int64_t n, double *out)
{
    if (n < 0 || (n > 0 && !out)) return fail(PT_E_INVALID, "pt_philox_fill: bad arguments");
    for (int64_t i = 0; i < n; ++i) out[i] = pt_philox_draw(seed, path, first_draw + (uint64_t)i);
    return PT_OK;
}

const char *pt_build_info(void)
{
    return "libptrans " "1" " gfx950 HIP " __DATE__ " fp64-exact";
}

size_t pt_last_error(char *buf, size_t len)
{
    if (buf && len) {
        size_t n = std::min(len - 1, g_err.size());
        memcpy(buf, g_err.data(), n);
        buf[n] = 0;
    }
    return g_err.size();
}

int pt_ctx_create(int device, uint64_t seed, pt_ctx **out)
{
    if (!out) return fail(PT_E_INVALID, "pt_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PT_E_NODEVICE, "pt_ctx_create: no HIP device");
    if (device < 0 || device >= ndev)
        return fail(PT_E_INVALID, "pt_ctx_create: device %d out of range (%d devices)", device, ndev);
    PT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PT_HIP(hipGetDeviceProperties(&prop, device));
    pt_ctx *c = new pt_ctx();
    c->device = device;
    c->seed = seed;
    c->ncu = prop.multiProcessorCount;
    c->lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                      : prop.sharedMemPerMultiprocessor;
    int wall_khz = 0;
    PT_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));
    c->live_ticks = (int64_t)(wall_khz > 0 ? wall_khz : 100000) * 20000;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreate(&c->ev2);
    if (e == hipSuccess) e = hipMalloc((void **)&c->counters, kCounterWords * sizeof(unsigned long long));
    if (e == hipSuccess) e = scratch_address(device, &c->noalloc);
    if (e != hipSuccess) {
        pt_ctx_destroy(c);
        return fail(PT_E_HIP, "pt_ctx_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return PT_OK;
}

int pt_ctx_destroy(pt_ctx *c)
{
    if (!c) return PT_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->outcome) (void)hipFree(c->outcome);
    if (c->grid_buf) (void)hipFree(c->grid_buf);
    if (c->host_stage) (void)hipHostFree(c->host_stage);
    if (c->counters) (void)hipFree(c->counters);
    if (c->scale) (void)hipFree(c->scale);
    if (c->work) (void)hipFree(c->work);
    if (c->tail) (void)hipFree(c->tail);
    if (c->tail_score) (void)hipFree(c->tail_score);
    if (c->first_ids) (void)hipFree(c->first_ids);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PT_OK;
}

int pt_ctx_set_seed(pt_ctx *c, uint64_t seed)
{
    if (!c) return fail(PT_E_INVALID, "pt_ctx_set_seed: NULL context");
    c->seed = seed;
    return PT_OK;
}

int pt_ctx_set_event_limit(pt_ctx *c, int64_t limit)
{
    if (!c || limit < 1) return fail(PT_E_INVALID, "pt_ctx_set_event_limit: bad argument");
    if (limit > kMaxEventLimit)
        return fail(PT_E_INVALID, "pt_ctx_set_event_limit: %lld events exceed the maximum %lld "
                                  "(a path's 32-bit draw counter must not wrap)",
                    (long long)limit, (long long)kMaxEventLimit);
    c->event_limit = limit;
    return PT_OK;
}

int pt_ctx_set_tuning(pt_ctx *c, int32_t chunk, int32_t refill_min, int32_t blocks_per_cu)
{
    if (!c) return fail(PT_E_INVALID, "pt_ctx_set_tuning: NULL context");
    if (chunk < 1 || refill_min < 1 || refill_min > 64 || blocks_per_cu < 0)
        return fail(PT_E_INVALID, "pt_ctx_set_tuning: chunk >= 1, 1 <= refill_min <= 64, "
                                  "blocks_per_cu >= 0 (0: occupancy maximum)");
    c->chunk = chunk;
    c->refill_min = refill_min;
    c->blocks_per_cu = blocks_per_cu;
    return PT_OK;
}

int pt_ctx_set_tail(pt_ctx *c, int64_t events, int64_t max_records, int32_t live_blocks)
{
    if (!c) return fail(PT_E_INVALID, "pt_ctx_set_tail: NULL context");
    if (events < 1 || max_records < 0 || max_records > ((int64_t)1 << 26) || live_blocks < 0 ||
        live_blocks > 64)
        return fail(PT_E_INVALID, "pt_ctx_set_tail: events >= 1, 0 <= max_records <= 2^26, "
                                  "0 <= live_blocks <= 64");
    c->tail_events = events;
    c->tail_max = max_records;
    c->live_blocks = live_blocks;
    return PT_OK;
}

int pt_ctx_set_op_counting(pt_ctx *c, int32_t on)
{
    if (!c) return fail(PT_E_INVALID, "pt_ctx_set_op_counting: NULL context");
    c->count_ops = on != 0;
    return PT_OK;
}

int pt_surface_setup_polygon(int32_t nv, const double *v, double center[3], double normal[3],
                             double tangent[9])
{
    // setupSurfaceGeometry, polygon branch (src/RTMCphon.f90:1440-1487)
    if (nv < 3 || !v || !center || !normal || !tangent)
        return fail(PT_E_INVALID, "pt_surface_setup_polygon: need >= 3 vertices");
    auto V = [&](int i) { return v + 3 * ((i - 1) % nv); };   // 1-based, closed loop
    double c[3] = {0.0, 0.0, 0.0};
    for (int iv = 1; iv <= nv; ++iv)
        for (int k = 0; k < 3; ++k) c[k] = c[k] + V(iv)[k];
    const double inv = 1.0 / nv;
    for (int k = 0; k < 3; ++k) center[k] = inv * c[k];
    double n[3] = {0.0, 0.0, 0.0};
    auto add_unit_cross = [&](const double *p0, const double *p1, const double *p2) {
        double a[3], b[3], cr[3];
        for (int k = 0; k < 3; ++k) { a[k] = p1[k] - p0[k]; b[k] = p2[k] - p0[k]; }
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        const double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int k = 0; k < 3; ++k) n[k] = n[k] + cr[k] / len;
    };
    for (int iv = 1; iv <= nv - 1; ++iv) add_unit_cross(V(iv), V(iv + 1), V(iv + 2));
    add_unit_cross(V(nv), V(nv + 1), V(2));
    for (int k = 0; k < 3; ++k) normal[k] = inv * n[k];
    tangent[6] = normal[0]; tangent[7] = normal[1]; tangent[8] = normal[2];
    if (std::fabs(normal[2]) == 1.0 ||
        (normal[0] * normal[0] == 0.0 && normal[1] * normal[1] == 0.0)) {
        tangent[3] = 0.0; tangent[4] = 1.0; tangent[5] = 0.0;
        tangent[0] = 1.0; tangent[1] = 0.0; tangent[2] = 0.0;
    } else {
        const double r = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1]);
        tangent[3] = -normal[1] / r;
        tangent[4] = normal[0] / r;
        tangent[5] = 0.0;
        const double *a = tangent + 3, *b = tangent + 6;
        tangent[0] = a[1] * b[2] - a[2] * b[1];
        tangent[1] = a[2] * b[0] - a[0] * b[2];
        tangent[2] = a[0] * b[1] - a[1] * b[0];
    }
    return PT_OK;
}

// True when grain 0 is exactly the six walls setupBoxGrain builds (RT:1266-1331): order
// x-, x+, y-, y+, z-, z+, internal quads on the box faces with axis-aligned unit normals.
static bool box_is_axis_aligned(const pt_geom_desc *d)
{
    if (d->grain_offset[1] != 6) return false;
    for (int w = 0; w < 6; ++w) {
        const pt_surface &s = d->surfaces[w];
        const int a = w / 2;
        const double P = (w & 1) ? d->box[a] : 0.0;
        if (s.kind != PT_SHAPE_POLYGON || s.nvertex != 4 || !s.internal) return false;
        for (int k = 0; k < 3; ++k) {
            const double want = k == a ? ((w & 1) ? 1.0 : -1.0) : 0.0;
            if (s.normal[k] != want) return false;
        }
        for (int i = 0; i < 4; ++i) {
            const double *v = d->vertices + 3 * (s.vertex_offset + i);
            if (v[a] != P) return false;
            for (int k = 0; k < 3; ++k)
                if (v[k] != 0.0 && v[k] != d->box[k]) return false;
        }
    }
    return true;
}

int pt_geom_create(pt_ctx *c, const pt_geom_desc *d, pt_geom **out)
{
    if (!c || !d || !out) return fail(PT_E_INVALID, "pt_geom_create: NULL argument");
    *out = nullptr;
    const int ng = d->ngrain_total, ns = d->nsurface;
    if (d->kind_grain != PT_GRAIN_VORONOI && d->kind_grain != PT_GRAIN_BRICKS &&
        d->kind_grain != PT_GRAIN_MESH)
        return fail(PT_E_GEOMETRY, "kind_grain %d is not voronoi, bricks or mesh", d->kind_grain);
    const int expect = d->kind_grain == PT_GRAIN_BRICKS ? d->ngrain + 2 : d->ngrain + 1;
    if (d->ngrain < 0 || ng != expect)
        return fail(PT_E_GEOMETRY, "ngrain_total %d, expected %d for ngrain %d", ng, expect, d->ngrain);
    if (d->kind_grain == PT_GRAIN_BRICKS && d->igrain_fill != d->ngrain + 1)
        return fail(PT_E_GEOMETRY, "igrain_fill must be ngrain+1 for bricks");
    if (!d->grain_offset || !d->surfaces || (ns > 0 && d->nvertex > 0 && !d->vertices) ||
        !d->grain_incident || !d->generatrix)
        return fail(PT_E_INVALID, "pt_geom_create: NULL array in the descriptor");
    if (ns >= (1 << 30))      // the trace flags candidate indices with bit 30
        return fail(PT_E_GEOMETRY, "nsurface %d: at most 2^30 - 1 surfaces", ns);
    if (d->grain_offset[0] != 0 || d->grain_offset[ng] != ns)
        return fail(PT_E_GEOMETRY, "grain_offset must run from 0 to nsurface");
    for (int g = 0; g < ng; ++g)
        if (d->grain_offset[g + 1] < d->grain_offset[g])
            return fail(PT_E_GEOMETRY, "grain_offset is not monotone at grain %d", g);
    if (d->grain_offset[1] < 2)
        return fail(PT_E_GEOMETRY, "the box grain needs its x- and x+ walls first");
    if (!(d->box[0] > 0 && d->box[1] > 0 && d->box[2] > 0))
        return fail(PT_E_GEOMETRY, "box dimensions must be positive");
    if (d->nvertex > INT32_MAX / 3) return fail(PT_E_GEOMETRY, "too many vertices");

    std::vector<DSurf> surf(ns);
    std::vector<double> tang(9 * (size_t)ns);
    for (int i = 0; i < ns; ++i) {
        const pt_surface &s = d->surfaces[i];
        DSurf &o = surf[i];
        memset(&o, 0, sizeof o);
        if (s.kind != PT_SHAPE_POLYGON && s.kind != PT_SHAPE_SPHERE && s.kind != PT_SHAPE_CYLINDER)
            return fail(PT_E_GEOMETRY, "surface %d: unknown shape %d", i, s.kind);
        if (s.kind == PT_SHAPE_POLYGON) {
            if (s.nvertex < 3 || s.vertex_offset < 0 || s.vertex_offset + s.nvertex > d->nvertex)
                return fail(PT_E_GEOMETRY, "surface %d: bad polygon vertex range", i);
        } else if (!(s.radius > 0)) {
            return fail(PT_E_GEOMETRY, "surface %d: radius must be positive", i);
        }
        if (s.internal) {
            if (s.trans_kind < PT_TRANS_VALUE || s.trans_kind > PT_TRANS_PERIODIC)
                return fail(PT_E_GEOMETRY, "surface %d: unknown transmissivity law %d", i, s.trans_kind);
            if (s.trans_kind == PT_TRANS_LORENTZ && !(s.trans_param > 0.0))
                return fail(PT_E_GEOMETRY, "surface %d: Lorentzian omega0 must be positive", i);
            if (s.trans_kind == PT_TRANS_PERIODIC && (i < 2 || i >= 6 || d->grain_offset[1] != 6))
                return fail(PT_E_GEOMETRY, "surface %d: only box side walls can be periodic", i);
            if (s.grain_outside < 0 || s.grain_outside >= ng)
                return fail(PT_E_GEOMETRY, "surface %d: grain_outside %d out of range", i, s.grain_outside);
        }
        for (int k = 0; k < 3; ++k) { o.n[k] = s.normal[k]; o.c[k] = s.center[k]; }
        o.r = s.radius;
        o.p = s.specularity;
        o.tparam = s.trans_param;
        o.kind = s.kind;
        o.nv = s.kind == PT_SHAPE_POLYGON ? s.nvertex : 0;
        o.voff = (int32_t)s.vertex_offset;
        o.gout = s.grain_outside;
        o.tkind = s.trans_kind;
        o.gin = s.grain_inside;
        memcpy(&tang[9 * (size_t)i], s.tangent, 9 * sizeof(double));
    }
    // box walls x- and x+ must be the first two surfaces (EndSide, RT:1748-1753)
    std::vector<int32_t> cand_off(ng + 1, 0), cand_idx, inc_off(ng + 1, 0), inc_idx;
    for (int g = 0; g < ng; ++g) {
        for (int i = d->grain_offset[g]; i < d->grain_offset[g + 1]; ++i) {
            if (d->surfaces[i].internal) cand_idx.push_back(i);
            if (d->surfaces[i].incident && d->surfaces[i].kind == PT_SHAPE_POLYGON) inc_idx.push_back(i);
        }
        cand_off[g + 1] = (int32_t)cand_idx.size();
        inc_off[g + 1] = (int32_t)inc_idx.size();
    }
    // per-lane scans (Voronoi-type grains: only polygon candidates, in an aligned box): a
    // plane record per candidate
    bool lane_scan = box_is_axis_aligned(d) && d->kind_grain == PT_GRAIN_VORONOI;
    for (size_t j = (size_t)cand_off[1]; j < cand_idx.size() && lane_scan; ++j)
        lane_scan = d->surfaces[cand_idx[j]].kind == PT_SHAPE_POLYGON;
    lane_scan = lane_scan && cand_idx.size() > (size_t)cand_off[1];
    std::vector<CandPlane> cplane;
    if (lane_scan) {
        cplane.resize(cand_idx.size());
        for (size_t j = 0; j < cand_idx.size(); ++j) {
            const pt_surface &s = d->surfaces[cand_idx[j]];
            CandPlane &q = cplane[j];
            memset(&q, 0, sizeof q);
            for (int k = 0; k < 3; ++k) {
                q.n[k] = s.normal[k];
                q.nf[k] = (float)s.normal[k];
            }
            if (s.kind == PT_SHAPE_POLYGON)
                for (int k = 0; k < 3; ++k) q.v1[k] = d->vertices[3 * s.vertex_offset + k];
            q.idx = cand_idx[j];
        }
    }
    // regions of the tail kernel (lane_scan geometries): per grain the planes of its scan
    // order, normalised and oriented towards the grain's generatrix
    std::vector<float4> region;
    std::vector<int32_t> region_off(lane_scan ? ng + 1 : 0, 0);
    if (lane_scan) {
        for (int gi = 0; gi < ng; ++gi) {
            region_off[gi] = (int32_t)region.size();
            if (gi == 0) continue;
            const double *gp = d->generatrix + 3 * (size_t)gi;
            const size_t mark = region.size();
            bool ok = true;
            for (int part = 0; part < 2 && ok; ++part) {
                const int g2 = part == 0 ? 0 : gi;
                for (int j = cand_off[g2]; j < cand_off[g2 + 1] && ok; ++j) {
                    const pt_surface &s = d->surfaces[cand_idx[j]];
                    const double *v = d->vertices + 3 * s.vertex_offset;
                    const double nn = std::sqrt(s.normal[0] * s.normal[0] + s.normal[1] * s.normal[1] +
                                                s.normal[2] * s.normal[2]);
                    if (!(nn > 0.0)) { ok = false; break; }
                    double u[3], off = 0.0, side = 0.0;
                    for (int k = 0; k < 3; ++k) u[k] = s.normal[k] / nn;
                    for (int k = 0; k < 3; ++k) { off += u[k] * v[k]; side += u[k] * (gp[k] - v[k]); }
                    if (!(std::fabs(side) > 1.0e-3)) { ok = false; break; }
                    const double sg = side > 0.0 ? 1.0 : -1.0;
                    region.push_back(make_float4((float)(sg * u[0]), (float)(sg * u[1]), (float)(sg * u[2]),
                                                 (float)(sg * off)));
                }
            }
            if (!ok) region.resize(mark);
        }
        region_off[ng] = (int32_t)region.size();
    }
    // start-grain map: a cell of the start face (y, z at x = 0) gets grain g when, on the
    // whole cell widened by a rounding margin, both start-grain searches certainly return
    // g.  Voronoi: on all four corners g is the nearest incident generatrix by 1e-4 nm^2
    // (the differences of squared distances are affine in p).  Bricks whose grains are
    // none incident: the fill grain.  Then g's candidates are all certainly inside and
    // every other grain has a candidate certainly outside (polygons: (p - c).n, affine,
    // beyond m = 1e-6 |n| on all corners; spheres and z-cylinders: the squared distance to
    // the centre (axis) over the cell, at least the box distance and at most the farthest
    // corner's, beyond r^2 -+ 1e-6 (r^2 + 1)).  The margins are far beyond the FP64
    // rounding of the device's tests.
    std::vector<int32_t> start_map;
    const int kStartGrid = 64;
    double start_h[2] = {d->box[1] * d->scope[0] / kStartGrid, d->box[2] * d->scope[1] / kStartGrid};
    bool any_incident = false;
    for (int gi = 1; gi <= d->ngrain; ++gi) any_incident = any_incident || d->grain_incident[gi];
    const bool map_voronoi = d->kind_grain == PT_GRAIN_VORONOI;
    bool map_ok = (map_voronoi || (d->kind_grain == PT_GRAIN_BRICKS && !any_incident &&
                                   d->igrain_fill == d->ngrain + 1)) &&
                  d->ngrain > 0 && start_h[0] > 0.0 && start_h[1] > 0.0;
    for (int gi = 1; gi <= d->ngrain && map_ok; ++gi)
        for (int j = cand_off[gi]; j < cand_off[gi + 1] && map_ok; ++j) {
            const int kd = d->surfaces[cand_idx[j]].kind;
            map_ok = kd == PT_SHAPE_POLYGON || (!map_voronoi && (kd == PT_SHAPE_SPHERE || kd == PT_SHAPE_CYLINDER));
        }
    if (map_ok) {
        start_map.assign((size_t)kStartGrid * kStartGrid, 0);
        auto d2 = [&](int gi, const double *p) {
            const double *q = d->generatrix + 3 * (size_t)gi;
            return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
        };
        // +1: candidate j certainly inside (inside_surface true) on the cell, -1: certainly
        // outside, 0: unknown
        auto side = [&](int j, double cs[4][3], const double *lo, const double *hi) {
            const pt_surface &sf = d->surfaces[cand_idx[j]];
            if (sf.kind == PT_SHAPE_POLYGON) {
                const double m = 1.0e-6 * std::sqrt(sf.normal[0] * sf.normal[0] + sf.normal[1] * sf.normal[1] +
                                                    sf.normal[2] * sf.normal[2]);
                double vmin = 1e300, vmax = -1e300;
                for (int c4 = 0; c4 < 4; ++c4) {
                    double v = 0.0;
                    for (int k = 0; k < 3; ++k) v += (cs[c4][k] - sf.center[k]) * sf.normal[k];
                    vmin = std::min(vmin, v);
                    vmax = std::max(vmax, v);
                }
                return vmax < -m ? 1 : (vmin > m ? -1 : 0);
            }
            const int nk = sf.kind == PT_SHAPE_SPHERE ? 3 : 2;
            double qmin = 0.0, qmax = 0.0;
            for (int k = 0; k < nk; ++k) {
                const double a = lo[k] - sf.center[k], b = hi[k] - sf.center[k];
                const double near = (a <= 0.0 && b >= 0.0) ? 0.0 : std::min(std::fabs(a), std::fabs(b));
                const double far = std::max(std::fabs(a), std::fabs(b));
                qmin += near * near;
                qmax += far * far;
            }
            const double r2 = sf.radius * sf.radius, m = 1.0e-6 * (r2 + 1.0);
            return qmax < r2 - m ? 1 : (qmin > r2 + m ? -1 : 0);
        };
        for (int iy = 0; iy < kStartGrid; ++iy)
            for (int iz = 0; iz < kStartGrid; ++iz) {
                const double ey = 1.0e-6 * start_h[0] + 1.0e-9, ez = 1.0e-6 * start_h[1] + 1.0e-9;
                const double lo[3] = {0.0, iy * start_h[0] - ey, iz * start_h[1] - ez};
                const double hi[3] = {0.0, (iy + 1) * start_h[0] + ey, (iz + 1) * start_h[1] + ez};
                double cs[4][3];
                for (int c4 = 0; c4 < 4; ++c4) {
                    cs[c4][0] = 0.0;
                    cs[c4][1] = (c4 & 1) ? hi[1] : lo[1];
                    cs[c4][2] = (c4 & 2) ? hi[2] : lo[2];
                }
                int best = 0;
                bool ok = true;
                if (map_voronoi) {
                    double bd = 1.0e10;
                    for (int gi = 1; gi <= d->ngrain; ++gi)
                        if (d->grain_incident[gi] && bd > d2(gi, cs[0])) { bd = d2(gi, cs[0]); best = gi; }
                    ok = best > 0;
                    for (int gi = 1; gi <= d->ngrain && ok; ++gi)
                        if (gi != best && d->grain_incident[gi])
                            for (int c4 = 0; c4 < 4 && ok; ++c4) ok = d2(gi, cs[c4]) - d2(best, cs[c4]) > 1.0e-4;
                } else {
                    best = d->igrain_fill;
                }
                for (int gi = 1; gi <= d->ngrain && ok; ++gi) {
                    if (gi == best) {
                        for (int j = cand_off[gi]; j < cand_off[gi + 1] && ok; ++j) ok = side(j, cs, lo, hi) == 1;
                    } else {
                        bool out = false;
                        for (int j = cand_off[gi]; j < cand_off[gi + 1] && !out; ++j) out = side(j, cs, lo, hi) == -1;
                        ok = out;
                    }
                }
                start_map[(size_t)iy * kStartGrid + iz] = ok ? best : 0;
            }
    }
    std::vector<int32_t> ginc(d->grain_incident, d->grain_incident + ng);
    std::vector<double> gen(d->generatrix, d->generatrix + 3 * (size_t)ng);
    std::vector<double> vtx(d->vertices, d->vertices + 3 * (size_t)d->nvertex);

    PT_HIP(hipSetDevice(c->device));
    pt_geom *G = new pt_geom();
    G->ctx = c;
    int rc = PT_OK;
    DSurf *d_surf = nullptr;
    double *d_tang = nullptr, *d_vtx = nullptr, *d_gen = nullptr;
    int32_t *d_co = nullptr, *d_ci = nullptr, *d_io = nullptr, *d_ii = nullptr, *d_ginc = nullptr;
    CandPlane *d_cp = nullptr;
    float4 *d_rg = nullptr;
    int32_t *d_ro = nullptr, *d_sm = nullptr;
    if ((rc = upload(surf, &d_surf, c->stream)) || (rc = upload(tang, &d_tang, c->stream)) ||
        (rc = upload(vtx, &d_vtx, c->stream)) || (rc = upload(gen, &d_gen, c->stream)) ||
        (rc = upload(cand_off, &d_co, c->stream)) || (rc = upload(cand_idx, &d_ci, c->stream)) ||
        (rc = upload(inc_off, &d_io, c->stream)) || (rc = upload(inc_idx, &d_ii, c->stream)) ||
        (rc = upload(ginc, &d_ginc, c->stream)) || (rc = upload(cplane, &d_cp, c->stream)) ||
        (rc = upload(region, &d_rg, c->stream)) || (rc = upload(region_off, &d_ro, c->stream)) ||
        (rc = upload(start_map, &d_sm, c->stream))) {
        for (void *p : {(void *)d_surf, (void *)d_tang, (void *)d_vtx, (void *)d_gen, (void *)d_co,
                        (void *)d_ci, (void *)d_io, (void *)d_ii, (void *)d_ginc, (void *)d_cp,
                        (void *)d_rg, (void *)d_ro, (void *)d_sm})
            if (p) (void)hipFree(p);
        delete G;
        return rc;
    }
    G->owned = {d_surf, d_tang, d_vtx, d_gen, d_co, d_ci, d_io, d_ii, d_ginc, d_cp, d_rg, d_ro, d_sm};
    G->surf = surf;
    DevGeom &g = G->dg;
    g.surf = d_surf;
    g.tang = d_tang;
    g.vtx = d_vtx;
    g.cand_off = d_co;
    g.cand_idx = d_ci;
    g.inc_off = d_io;
    g.inc_idx = d_ii;
    g.grain_incident = d_ginc;
    g.gen = d_gen;
    g.surf_base = nullptr;
    g.cand_plane = d_cp;
    g.lane_scan = lane_scan ? 1 : 0;
    g.region = d_rg;
    g.region_off = lane_scan ? d_ro : nullptr;
    // FP32 region tests: coordinates below the box's largest side, a few roundings of
    // 2^-24 of it each; the margin is far above them and the reference's FP64 rounding
    g.region_margin = (float)(1.0e-2 + 2.0e-5 * std::max(d->box[0], std::max(d->box[1], d->box[2])));
    g.start_map = d_sm;
    g.start_grid = d_sm ? kStartGrid : 0;
    // the cell of a start point from a product (a rounding away from the quotient; the
    // certified cells are widened by 1e-6 of a side)
    g.start_ih[0] = start_h[0] > 0.0 ? 1.0 / start_h[0] : 0.0;
    g.start_ih[1] = start_h[1] > 0.0 ? 1.0 / start_h[1] : 0.0;
    for (int k = 0; k < 3; ++k) g.box[k] = d->box[k];
    g.scope[0] = d->scope[0];
    g.scope[1] = d->scope[1];
    g.ngrain = d->ngrain;
    g.kind_grain = d->kind_grain;
    g.igrain_fill = d->igrain_fill;
    g.nmaterial = d->nmaterial;
    g.ngrain_total = ng;
    g.nsurface = ns;
    g.box_aligned = box_is_axis_aligned(d) ? 1 : 0;
    g.periodic = 0;
    for (int a = 1; a < 3 && ns >= 6; ++a) {
        const bool lo = d->surfaces[2 * a].trans_kind == PT_TRANS_PERIODIC;
        const bool hi = d->surfaces[2 * a + 1].trans_kind == PT_TRANS_PERIODIC;
        if (lo != hi) rc = fail(PT_E_GEOMETRY, "periodic walls come in pairs (axis %d)", a);
        if (!lo) continue;
        g.periodic = 1;
        if (!g.box_aligned) rc = fail(PT_E_GEOMETRY, "periodic walls need the axis-aligned box");
        if (d->kind_grain == PT_GRAIN_VORONOI && d->ngrain > 1)
            rc = fail(PT_E_GEOMETRY, "periodic walls need a periodic tessellation: not with "
                                     "several Voronoi grains");
        for (int i = 6; i < ns && rc == PT_OK; ++i) {
            const pt_surface &s = d->surfaces[i];
            if (!s.internal) continue;
            if (s.kind == PT_SHAPE_SPHERE &&
                (s.center[a] - s.radius < 0.0 || s.center[a] + s.radius > d->box[a]))
                rc = fail(PT_E_GEOMETRY, "sphere %d crosses a periodic wall", i);
            if (s.kind == PT_SHAPE_CYLINDER && a == 1 &&
                (s.center[1] - s.radius < 0.0 || s.center[1] + s.radius > d->box[1]))
                rc = fail(PT_E_GEOMETRY, "cylinder %d crosses a periodic wall", i);
            if (s.kind == PT_SHAPE_POLYGON)
                for (int k = 0; k < s.nvertex; ++k) {
                    const double v = d->vertices[3 * (s.vertex_offset + k) + a];
                    if (v < 0.0 || v > d->box[a])
                        rc = fail(PT_E_GEOMETRY, "polygon %d crosses a periodic wall", i);
                }
        }
    }
    if (rc != PT_OK) {
        pt_geom_destroy(G);
        return rc;
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        pt_geom_destroy(G);
        return fail(PT_E_HIP, "pt_geom_create: %s", hipGetErrorString(e));
    }
    *out = G;
    return PT_OK;
}

int pt_geom_set_specularity(pt_geom *G, const double *p)
{
    if (!G || !p) return fail(PT_E_INVALID, "pt_geom_set_specularity: NULL argument");
    const int ns = G->dg.nsurface;
    for (int i = 0; i < ns; ++i)
        if (!(p[i] >= 0.0 && p[i] <= 1.0))
            return fail(PT_E_INVALID, "pt_geom_set_specularity: specularity %d is %g, outside [0, 1]", i, p[i]);
    for (int i = 0; i < ns; ++i) G->surf[i].p = p[i];
    PT_HIP(hipSetDevice(G->ctx->device));
    hipError_t e = hipMemcpyAsync((void *)G->dg.surf, G->surf.data(), ns * sizeof(DSurf),
                                  hipMemcpyHostToDevice, G->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G->ctx->stream);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_geom_set_specularity: %s", hipGetErrorString(e));
    return PT_OK;
}

int pt_geom_destroy(pt_geom *G)
{
    if (!G) return PT_OK;
    (void)hipSetDevice(G->ctx->device);
    (void)hipStreamSynchronize(G->ctx->stream);
    for (void *p : G->owned)
        if (p) (void)hipFree(p);
    delete G;
    return PT_OK;
}

int pt_alias_build(int64_t n, const double *w, double *prob, int32_t *alias)
{
    // Vose's alias method; small and large work lists are processed from their ends.
    if (n < 1 || n > INT32_MAX || !w || !prob || !alias)
        return fail(PT_E_INVALID, "pt_alias_build: bad arguments");
    double tot = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(w[i] >= 0.0) || !std::isfinite(w[i]))
            return fail(PT_E_INVALID, "pt_alias_build: weight %lld is negative or not finite", (long long)i);
        tot += w[i];
    }
    if (!(tot > 0.0)) return fail(PT_E_INVALID, "pt_alias_build: all weights are zero");
    std::vector<double> q(n);
    std::vector<int64_t> sm, lg;
    sm.reserve(n);
    lg.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        q[i] = w[i] * (double)n / tot;
        alias[i] = (int32_t)i;
        if (q[i] < 1.0) sm.push_back(i); else lg.push_back(i);
    }
    while (!sm.empty() && !lg.empty()) {
        const int64_t a = sm.back(), b = lg.back();
        sm.pop_back();
        lg.pop_back();
        prob[a] = q[a];
        alias[a] = (int32_t)b;
        q[b] = (q[b] + q[a]) - 1.0;
        if (q[b] < 1.0) sm.push_back(b); else lg.push_back(b);
    }
    while (!lg.empty()) { prob[lg.back()] = 1.0; lg.pop_back(); }
    while (!sm.empty()) { prob[sm.back()] = 1.0; sm.pop_back(); }
    return PT_OK;
}

int pt_modes_create(pt_ctx *c, const pt_modes_desc *d, pt_modes **out)
{
    if (!c || !d || !out) return fail(PT_E_INVALID, "pt_modes_create: NULL argument");
    *out = nullptr;
    if (d->nbranch < 1 || d->nkpoint < 1) return fail(PT_E_INVALID, "pt_modes_create: empty table");
    if (!d->omega) return fail(PT_E_INVALID, "pt_modes_create: omega is NULL");
    const size_t n = (size_t)d->nbranch * d->nkpoint;
    if (n > (size_t)INT32_MAX) return fail(PT_E_INVALID, "pt_modes_create: table too large");
    std::vector<double> om(d->omega, d->omega + n);
    const bool alias = d->vel && d->mfp && d->weight;
    std::vector<double> mfp, prob(n), wts(n);
    std::vector<int32_t> al(n);
    if (alias) {
        if (!(d->temperature > 0)) return fail(PT_E_INVALID, "pt_modes_create: temperature must be positive");
        mfp.assign(d->mfp, d->mfp + n);
        // heat capacity, src/RTMCphon.f90:2696-2701
        const double hbar = 1.054571e-34, kb = 1.3806503e-23;
        for (size_t i = 0; i < n; ++i) {
            const double x = hbar * om[i] / (kb * d->temperature);
            const double cap = x <= 1.0e-10 ? kb : kb * std::pow((x / 2) / std::sinh(x / 2), 2.0);
            wts[i] = cap * d->vel[i] * d->mfp[i] * d->weight[i];
            if (!(mfp[i] > 0)) return fail(PT_E_INVALID, "pt_modes_create: mfp[%zu] must be positive", i);
        }
        int rc = pt_alias_build((int64_t)n, wts.data(), prob.data(), al.data());
        if (rc) return rc;
    }
    PT_HIP(hipSetDevice(c->device));
    pt_modes *M = new pt_modes();
    M->ctx = c;
    M->nb = d->nbranch;
    M->nk = d->nkpoint;
    M->has_alias = alias;
    double *d_om = nullptr, *d_mfp = nullptr, *d_imfp = nullptr, *d_prob = nullptr;
    int32_t *d_al = nullptr;
    // 1 / mfp on the host: IEEE division rounds as the device's would
    std::vector<double> imfp(mfp.size());
    for (size_t i = 0; i < mfp.size(); ++i) imfp[i] = 1.0 / mfp[i];
    int rc = upload(om, &d_om, c->stream);
    if (!rc && alias) rc = upload(mfp, &d_mfp, c->stream);
    if (!rc && alias) rc = upload(imfp, &d_imfp, c->stream);
    if (!rc && alias) rc = upload(prob, &d_prob, c->stream);
    if (!rc && alias) rc = upload(al, &d_al, c->stream);
    M->owned = {d_om, d_mfp, d_imfp, d_prob, d_al};
    if (rc) { pt_modes_destroy(M); return rc; }
    M->dm.omega = d_om;
    M->dm.mfp = d_mfp;
    M->dm.inv_mfp = d_imfp;
    M->dm.prob = d_prob;
    M->dm.alias = d_al;
    M->dm.scale = nullptr;
    M->dm.nb = d->nbranch;
    M->dm.nk = d->nkpoint;
    M->dm.max_omega = d->max_omega;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        pt_modes_destroy(M);
        return fail(PT_E_HIP, "pt_modes_create: %s", hipGetErrorString(e));
    }
    *out = M;
    return PT_OK;
}

int pt_modes_destroy(pt_modes *M)
{
    if (!M) return PT_OK;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    for (void *p : M->owned)
        if (p) (void)hipFree(p);
    delete M;
    return PT_OK;
}

}  // extern "C"

namespace {

// param%vangle (RT:1554-1556) and the cos/sin of each angle with the trace's own
// sincos, so host and device see the same direction cosines.
void angle_tables(int32_t nangle, const double *vangle, std::vector<double> &ct,
                  std::vector<double> &st)
{
    ct.resize(nangle);
    st.resize(nangle);
    for (int ia = 1; ia <= nangle; ++ia) {
        const double th = vangle ? vangle[ia - 1]
                                 : 0.0 + (double)(ia - 1) / (double)(nangle - 1) * (kPi * 0.5);
        pt_sincos(th, &st[ia - 1], &ct[ia - 1]);
    }
}

int check_trace_args(pt_ctx *c, const pt_geom *G, const pt_modes *M, double mfp, int32_t nangle,
                     int64_t np)
{
    if (!c || !G || !M) return fail(PT_E_INVALID, "NULL context, geometry or modes");
    if (G->ctx != c || M->ctx != c) return fail(PT_E_INVALID, "handles belong to another context");
    if (nangle < 2) return fail(PT_E_INVALID, "nangle must be >= 2");
    if (np < 0) return fail(PT_E_INVALID, "np must be >= 0");
    if (!(mfp > 0)) return fail(PT_E_INVALID, "mean free path must be positive");
    // the kernels number cells in 32 bits and divide path ids in FP64
    const int64_t ncell = (int64_t)M->nb * M->nk * (nangle - 1);
    if (ncell >= ((int64_t)1 << 31)) return fail(PT_E_UNSUPPORTED, "more than 2^31 cells");
    if (np > 0 && ncell > (((int64_t)1 << 53) - 1) / np)
        return fail(PT_E_UNSUPPORTED, "more than 2^53 paths in one call");
    return PT_OK;
}

}  // namespace

namespace {

TraceParams base_params(pt_ctx *c, const pt_geom *G, const pt_modes *M)
{
    TraceParams P{};
    P.g = G->dg;
    P.m = M->dm;
    P.seed = c->seed;
    P.chunk = c->chunk;
    P.refill_min = c->refill_min;
    P.event_limit = c->event_limit;
    P.precision = PT_PREC_FP64;
    P.queue = c->counters;
    P.stats = c->counters + 1;
    // the op-counting kernels run both start-grain searches, as the reference does
    if (c->count_ops) P.g.start_map = nullptr;
    return P;
}

// The context's live tail kernel for launch_trace (nullptr: none)
const LiveTail *live_tail(pt_ctx *c, LiveTail *lt)
{
    if (c->live_blocks <= 0 || !c->side) return nullptr;
    lt->side = c->side;
    lt->fork = c->fork;
    lt->join = c->join;
    lt->blocks = c->live_blocks;
    lt->ticks = c->live_ticks;
    lt->ncu = c->ncu;
    lt->lds_cu = c->lds_cu;
    return lt;
}

// The tail buffer of a launch of `total` paths of (mode, precision) with ngroup adjoint
// scores per path: room for total / 256 records (at least 65536, at most the context's
// tail_max), grown as needed and kept.  No records (paths stay in the persistent kernel)
// for traces that hand nothing over or when tail_max is 0.
int tail_prepare(pt_ctx *c, int mode, int precision, int64_t total, int ngroup, TraceParams &P)
{
    P.tail = nullptr;
    P.tail_score = nullptr;
    P.tail_cap = 0;
    P.tail_events = c->tail_events;
    P.first_ids = nullptr;
    P.first_cap = 0;
    if (!tail_mode(mode, precision) || c->count_ops || c->tail_max <= 0) return PT_OK;
    if (mode == TRACE_ALIAS || mode == TRACE_ADJOINT || mode == TRACE_GRAINS || mode == TRACE_BATCH) {
        // the list covers the last quarter of the ids (batch: all of them, its queue runs the
        // structures in no order the ids tell): room for total / 1024 paths (at least 65536,
        // at most 2^24)
        const int64_t fcap = std::min<int64_t>((int64_t)1 << 24, std::max<int64_t>(65536, total / 1024));
        if (c->first_alloc < fcap) {
            if (c->first_ids) PT_HIP(hipFree(c->first_ids));
            c->first_ids = nullptr;
            c->first_alloc = 0;
            PT_HIP(hipMalloc((void **)&c->first_ids, (size_t)fcap * sizeof(int64_t)));
            c->first_alloc = fcap;
        }
        P.first_ids = c->first_ids;
        P.first_cap = fcap;
        P.first_from = mode == TRACE_BATCH ? 0 : total - total / 4;
        P.first_mfp = kFirstMfpFrac * std::min(P.g.box[0], std::min(P.g.box[1], P.g.box[2]));
    }
    const int64_t cap = std::min<int64_t>(c->tail_max, std::max<int64_t>(65536, total / 256));
    if (c->tail_alloc < cap) {
        if (c->tail) PT_HIP(hipFree(c->tail));
        c->tail = nullptr;
        c->tail_alloc = 0;
        PT_HIP(hipMalloc((void **)&c->tail, (size_t)cap * sizeof(TailRec)));
        c->tail_alloc = cap;
        // epoch 0 marks a record no launch has written
        PT_HIP(hipMemsetAsync(c->tail, 0, (size_t)cap * sizeof(TailRec), c->stream));
        c->tail_epoch = 0;
    }
    if (ngroup > 0 && c->tail_score_alloc < cap * ngroup) {
        if (c->tail_score) PT_HIP(hipFree(c->tail_score));
        c->tail_score = nullptr;
        c->tail_score_alloc = 0;
        PT_HIP(hipMalloc((void **)&c->tail_score, (size_t)cap * ngroup * sizeof(double)));
        c->tail_score_alloc = cap * ngroup;
    }
    P.tail = c->tail;
    P.tail_score = ngroup > 0 ? c->tail_score : nullptr;
    P.tail_cap = cap;
    P.tail_cap_p = cap / 4;
    P.tail_prio_mfp = kTailPrioMfpFrac * std::min(P.g.box[0], std::min(P.g.box[1], P.g.box[2]));
    if (++c->tail_epoch == 0) {
        // 2^32 launches on one buffer: clear it before the numbers repeat
        PT_HIP(hipMemsetAsync(c->tail, 0, (size_t)c->tail_alloc * sizeof(TailRec), c->stream));
        c->tail_epoch = 1;
    }
    P.tail_epoch = c->tail_epoch;
    return PT_OK;
}

// Workgroups of a persistent trace launch: CUs x (the tuned count, or the occupancy
// maximum of this kernel variant with `lds` bytes of dynamic LDS).
hipError_t launch_blocks(const pt_ctx *c, int mode, int precision, size_t lds, int *blocks)
{
    int per_cu = c->blocks_per_cu;
    if (per_cu <= 0) {
        hipError_t e = trace_occupancy(mode, precision, c->threads, lds, &per_cu);
        if (e != hipSuccess) return e;
        if (per_cu < 1) per_cu = 1;
    }
    *blocks = c->ncu * per_cu;
    return hipSuccess;
}

// Paths a wave takes from the queue at once: the context's chunk, cut down for a launch too
// small to give every wave eight chunks, so that the queue still balances long paths over
// the waves.  Results do not depend on it (paths are keyed by id).
int32_t wave_chunk(const pt_ctx *c, int64_t total, int blocks)
{
    const int64_t waves = std::max<int64_t>(1, (int64_t)blocks * (c->threads / 64));
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(c->chunk, total / (8 * waves)));
}

int set_scale(pt_ctx *c, const pt_modes *M, const double *scale, TraceParams &P)
{
    P.m.scale = nullptr;
    if (!scale) return PT_OK;
    for (int b = 0; b < M->nb; ++b)
        if (!(scale[b] > 0) || !std::isfinite(scale[b]))
            return fail(PT_E_INVALID, "branch_scale[%d] must be positive and finite", b);
    if (c->scale_n < M->nb) {
        if (c->scale) PT_HIP(hipFree(c->scale));
        c->scale = nullptr;
        c->scale_n = 0;
        PT_HIP(hipMalloc((void **)&c->scale, M->nb * sizeof(double)));
        c->scale_n = M->nb;
    }
    PT_HIP(hipMemcpyAsync(c->scale, scale, M->nb * sizeof(double), hipMemcpyHostToDevice, c->stream));
    P.m.scale = c->scale;
    return PT_OK;
}

// The context's work space, grown to at least `bytes` (256-byte aligned pieces are carved
// from it by the caller).
int ctx_work(pt_ctx *c, size_t bytes, char **out)
{
    if (bytes > c->work_bytes) {
        PT_HIP(hipStreamSynchronize(c->stream));
        if (c->work) PT_HIP(hipFree(c->work));
        c->work = nullptr;
        c->work_bytes = 0;
        PT_HIP(hipMalloc(&c->work, bytes));
        c->work_bytes = bytes;
    }
    *out = (char *)c->work;
    return PT_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

void fill_stats(pt_stats *stats, const unsigned long long *h, float ms, float ms_red)
{
    if (!stats) return;
    stats->paths = (int64_t)h[1];
    stats->events = (int64_t)h[2];
    stats->draws = (int64_t)h[3];
    stats->lost = (int64_t)h[4];
    stats->start_mismatch = (int64_t)h[5];
    stats->skipped = (int64_t)h[7];
    stats->kernel_ms = ms;
    stats->reduce_ms = ms_red;
    stats->fp64_ops = (int64_t)h[8];
    stats->philox_blocks = (int64_t)h[9];
    stats->retests = (int64_t)h[11];
    stats->fallbacks = (int64_t)h[12];
    stats->bvh_nodes = (int64_t)h[13];
    stats->bvh_tests = (int64_t)h[14];
    stats->event_limited = (int64_t)h[6];
}

int replay_common(pt_ctx *c, TraceParams &P, int mode, int64_t first, int64_t n,
                  pt_path_record *records, int64_t stride = 1, double *points = nullptr,
                  int64_t max_points = 0, int64_t *npoints = nullptr)
{
    pt_path_record *d_rec = nullptr;
    double *d_pts = nullptr;
    int64_t *d_npts = nullptr;
    PT_HIP(hipMalloc((void **)&d_rec, n * sizeof(pt_path_record)));
    hipError_t e = hipSuccess;
    if (points) {
        if ((e = hipMalloc((void **)&d_pts, (size_t)n * max_points * 3 * sizeof(double))) == hipSuccess)
            e = hipMalloc((void **)&d_npts, n * sizeof(int64_t));
    }
    while (e == hipSuccess) {
        if ((e = launch_replay(P, mode, first, n, stride, d_rec, d_pts, max_points, d_npts, c->stream))) break;
        if (records && (e = hipMemcpyAsync(records, d_rec, n * sizeof(pt_path_record), hipMemcpyDeviceToHost, c->stream))) break;
        if (points) {
            if ((e = hipMemcpyAsync(points, d_pts, (size_t)n * max_points * 3 * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream))) break;
            if ((e = hipMemcpyAsync(npoints, d_npts, n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream))) break;
        }
        e = hipStreamSynchronize(c->stream);
        break;
    }
    (void)hipFree(d_rec);
    if (d_pts) (void)hipFree(d_pts);
    if (d_npts) (void)hipFree(d_npts);
    if (e != hipSuccess) return fail(PT_E_HIP, "replay: %s", hipGetErrorString(e));
    return PT_OK;
}

static int adjoint_common(pt_ctx *c, const pt_geom *G, const pt_modes *M, int64_t npaths,
                          int32_t precision, const double *branch_scale,
                          const int32_t *surface_group, int32_t ngroup, double baseline,
                          int32_t nbatch, int64_t *mode_paths, int64_t *mode_trans,
                          double *batch_sums, double *grain_tally, pt_stats *stats, const char *fn)
{
    if (!c || !G || !M) return fail(PT_E_INVALID, "NULL context, geometry or modes");
    if (G->ctx != c || M->ctx != c) return fail(PT_E_INVALID, "handles belong to another context");
    if (!M->has_alias) return fail(PT_E_INVALID, "modes were created without vel, mfp and weight");
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    if (npaths < 0) return fail(PT_E_INVALID, "npaths must be >= 0");
    if (ngroup < 0 || ngroup > kAdjMaxGroups)
        return fail(PT_E_UNSUPPORTED, "ngroup %d: at most %d specularity groups", ngroup, kAdjMaxGroups);
    if (ngroup > 0 && !surface_group) return fail(PT_E_INVALID, "surface_group is NULL");
    if (nbatch < 1 || nbatch > 4096) return fail(PT_E_INVALID, "nbatch must be in 1..4096");
    if (!std::isfinite(baseline)) return fail(PT_E_INVALID, "baseline is not finite");
    if (!mode_paths || !mode_trans)
        return fail(PT_E_INVALID, "tally arrays are NULL");
    const int ns = G->dg.nsurface;
    std::vector<int32_t> grp(ns, -1);
    for (int i = 0; i < ns && ngroup > 0; ++i) {
        const int gi = surface_group[i];
        if (gi < -1 || gi >= ngroup)
            return fail(PT_E_INVALID, "surface_group[%d] = %d outside -1..%d", i, gi, ngroup - 1);
        if (gi >= 0 && !(G->surf[i].p > 0.0 && G->surf[i].p < 1.0))
            return fail(PT_E_INVALID, "surface %d of group %d has specularity %g: the score of a "
                                      "specularity needs 0 < p < 1", i, gi, G->surf[i].p);
        grp[i] = gi;
    }
    if (stats) memset(stats, 0, sizeof *stats);
    if (npaths == 0) return PT_OK;
    const size_t nm = (size_t)M->nb * M->nk;
    PT_HIP(hipSetDevice(c->device));
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    int rc = set_scale(c, M, branch_scale, P);
    if (rc) return rc;
    if (branch_scale)
        for (int b = 0; b < M->nb; ++b)
            if (!(branch_scale[b] > 0.0)) return fail(PT_E_INVALID, "branch_scale[%d] must be > 0", b);
    P.total = npaths;
    P.ngroup = ngroup;
    P.nbatch = nbatch;
    P.baseline = baseline;
    P.ngt = grain_tally ? G->dg.ngrain_total : 0;
    // grain tallies (no scores, no batch sums) or the likelihood ratio: two kernel modes
    if (grain_tally && (ngroup > 0 || batch_sums))
        return fail(PT_E_UNSUPPORTED, "%s: grain tallies and adjoint scores in one launch", fn);
    const int amode = grain_tally ? TRACE_GRAINS : TRACE_ADJOINT;
    const size_t nsum = (size_t)nbatch * adj_stride(P);
    const size_t ntal = 2 * (size_t)P.ngt;
    const size_t lds = adj_lds_bytes(P, c->threads);
    hipDeviceProp_t prop;
    PT_HIP(hipGetDeviceProperties(&prop, c->device));
    if (lds + kTraceStaticLds > prop.sharedMemPerBlock)
        return fail(PT_E_UNSUPPORTED, "adjoint needs %zu bytes of LDS per workgroup (limit %zu): "
                                      "fewer groups or batches", lds + kTraceStaticLds,
                    (size_t)prop.sharedMemPerBlock);
    const size_t b_tal = align256(2 * nm * sizeof(unsigned long long));
    const size_t b_grp = align256((size_t)ns * sizeof(int32_t));
    const size_t b_sum = align256(nsum * sizeof(double));
    char *wk = nullptr;
    rc = ctx_work(c, b_tal + b_grp + b_sum + align256(ntal * sizeof(double)), &wk);
    if (rc) return rc;
    unsigned long long *d_tal = (unsigned long long *)wk;
    int32_t *d_grp = (int32_t *)(wk + b_tal);
    double *d_sum = (double *)(wk + b_tal + b_grp);
    double *d_gt = ntal ? (double *)(wk + b_tal + b_grp + b_sum) : nullptr;
    hipError_t e = hipSuccess;
    std::vector<double> h_gt(ntal);
    std::vector<unsigned long long> h_tal(2 * nm);
    std::vector<double> h_sum(nsum);
    unsigned long long host_ctr[kCounterWords] = {0};
    float ms = 0.f;
    int blocks = 0;
    do {
        if (e) break;
        P.mode_paths = d_tal;
        P.mode_trans = d_tal + nm;
        P.surf_group = d_grp;
        P.adj = d_sum;
        P.grain_tally = d_gt;
        if ((e = launch_blocks(c, amode, precision, lds, &blocks))) break;
        P.chunk = wave_chunk(c, P.total, blocks);
        if ((rc = tail_prepare(c, amode, precision, npaths, amode == TRACE_ADJOINT ? ngroup : 0, P))) return rc;
        if ((e = hipMemcpyAsync(d_grp, grp.data(), (size_t)ns * sizeof(int32_t), hipMemcpyHostToDevice, c->stream))) break;
        if ((e = hipMemsetAsync(c->counters, 0, kCounterWords * sizeof(unsigned long long), c->stream))) break;
        if ((e = hipMemsetAsync(d_tal, 0, 2 * nm * sizeof(unsigned long long), c->stream))) break;
        if ((e = hipMemsetAsync(d_sum, 0, nsum * sizeof(double), c->stream))) break;
        if (ntal && (e = hipMemsetAsync(d_gt, 0, ntal * sizeof(double), c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        LiveTail lt;
        if ((e = launch_trace(P, amode, blocks, c->threads, c->count_ops, lds, c->stream, live_tail(c, &lt)))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = hipMemcpyAsync(h_tal.data(), d_tal, 2 * nm * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(h_sum.data(), d_sum, nsum * sizeof(double), hipMemcpyDeviceToHost, c->stream))) break;
        if (ntal && (e = hipMemcpyAsync(h_gt.data(), d_gt, ntal * sizeof(double), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(host_ctr, c->counters, kCounterWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipStreamSynchronize(c->stream))) break;
        e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "%s: %s", fn, hipGetErrorString(e));
    for (size_t i = 0; i < nm; ++i) {
        mode_paths[i] += (int64_t)h_tal[i];
        mode_trans[i] += (int64_t)h_tal[nm + i];
    }
    if (batch_sums)
        for (size_t i = 0; i < nsum; ++i) batch_sums[i] += h_sum[i];
    for (size_t i = 0; i < ntal; ++i) grain_tally[i] += h_gt[i];
    fill_stats(stats, host_ctr, ms, 0.f);
    if ((int64_t)host_ctr[1] != npaths)
        return fail(PT_E_HIP, "%s: %lld of %lld paths completed",
                    fn, (long long)host_ctr[1], (long long)npaths);
    return PT_OK;
}


}  // namespace

extern "C" {

int pt_trace_grid(pt_ctx *c, const pt_geom *G, const pt_modes *M, double mfp, int32_t nangle,
                  const double *vangle, int64_t np, int32_t precision, int64_t *counts,
                  pt_stats *stats)
{
    int rc = check_trace_args(c, G, M, mfp, nangle, np);
    if (rc) return rc;
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    if (!counts) return fail(PT_E_INVALID, "counts is NULL");
    const int64_t nb = M->nb, nk = M->nk;
    const int64_t ncell = nb * nk * (nangle - 1);
    const int64_t total = ncell * np;
    const size_t ncount = (size_t)(nb * nk * nangle);
    memset(counts, 0, ncount * sizeof(int64_t));
    if (stats) memset(stats, 0, sizeof *stats);
    if (total == 0) return PT_OK;

    PT_HIP(hipSetDevice(c->device));
    std::vector<double> ct, st;
    angle_tables(nangle, vangle, ct, st);
    if ((size_t)total + 16 > c->outcome_bytes) {
        if (c->outcome) PT_HIP(hipFree(c->outcome));
        c->outcome = nullptr;
        c->outcome_bytes = 0;
        PT_HIP(hipMalloc((void **)&c->outcome, (size_t)total + 16));
        c->outcome_bytes = (size_t)total + 16;
    }
    const size_t gbytes = 2 * (size_t)nangle * sizeof(double) + ncount * sizeof(long long);
    if (gbytes > c->grid_buf_bytes) {
        if (c->grid_buf) PT_HIP(hipFree(c->grid_buf));
        c->grid_buf = nullptr;
        c->grid_buf_bytes = 0;
        PT_HIP(hipMalloc(&c->grid_buf, gbytes));
        c->grid_buf_bytes = gbytes;
    }
    double *d_ang = (double *)c->grid_buf;
    long long *d_counts = (long long *)(d_ang + 2 * nangle);
    if (ncount * sizeof(long long) > c->host_stage_bytes) {
        if (c->host_stage) PT_HIP(hipHostFree(c->host_stage));
        c->host_stage = nullptr;
        c->host_stage_bytes = 0;
        PT_HIP(hipHostMalloc(&c->host_stage, ncount * sizeof(long long), hipHostMallocDefault));
        c->host_stage_bytes = ncount * sizeof(long long);
    }
    hipError_t e = hipSuccess;
    std::vector<double> ang(ct);
    ang.insert(ang.end(), st.begin(), st.end());
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    P.cos_t = d_ang;
    P.sin_t = d_ang + nangle;
    P.mfp = mfp;
    P.np = np;
    P.total = total;
    P.nangle = nangle;
    P.outcome = c->outcome;
    unsigned long long host_ctr[kCounterWords] = {0};
    float ms = 0.f, ms_red = 0.f;
    int blocks = 0;
    // a diffusive sampling mean free path (one that gets safety balls) gets the kernel that
    // serves the last paths of each wave with the whole wave
    const double *bx = G->dg.box;
    const double lim = kBallMfpFrac * std::fmin(bx[0], std::fmin(bx[1], bx[2]));
    const int gmode = precision != PT_PREC_FP64 && mfp < lim ? TRACE_GRID_TAIL : TRACE_GRID;
    do {
        if ((e = launch_blocks(c, gmode, precision, 0, &blocks))) break;
        P.chunk = wave_chunk(c, P.total, blocks);
        if ((rc = tail_prepare(c, gmode, precision, total, 0, P))) return rc;
        if ((e = hipMemcpyAsync(d_ang, ang.data(), 2 * nangle * sizeof(double), hipMemcpyHostToDevice, c->stream))) break;
        if ((e = hipMemsetAsync(c->counters, 0, kCounterWords * sizeof(unsigned long long), c->stream))) break;
        if ((e = hipMemsetAsync(d_counts, 0, ncount * sizeof(long long), c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        LiveTail lt;
        if ((e = launch_trace(P, gmode, blocks, c->threads, c->count_ops, 0, c->stream, live_tail(c, &lt)))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = launch_reduce_grid(c->outcome, ncell, np, nb, nk, nangle, d_counts, c->counters + 7, c->stream))) break;
        if ((e = hipEventRecord(c->ev2, c->stream))) break;
        if ((e = hipMemcpyAsync(c->host_stage, d_counts, ncount * sizeof(long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(host_ctr, c->counters, kCounterWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipStreamSynchronize(c->stream))) break;
        if ((e = hipEventElapsedTime(&ms, c->ev0, c->ev1))) break;
        e = hipEventElapsedTime(&ms_red, c->ev1, c->ev2);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_trace_grid: %s", hipGetErrorString(e));
    memcpy(counts, c->host_stage, ncount * sizeof(long long));
    fill_stats(stats, host_ctr, ms, ms_red);
    if ((int64_t)host_ctr[1] != total)
        return fail(PT_E_HIP, "pt_trace_grid: %lld of %lld paths completed",
                    (long long)host_ctr[1], (long long)total);
    return PT_OK;
}

static int grid_replay(pt_ctx *c, const pt_geom *G, const pt_modes *M, double mfp, int32_t nangle,
                       const double *vangle, int64_t np, int64_t first_path, int64_t stride,
                       int64_t n, int32_t precision, pt_path_record *records, double *points,
                       int64_t max_points, int64_t *npoints, const char *fn)
{
    int rc = check_trace_args(c, G, M, mfp, nangle, np);
    if (rc) return rc;
    if (np < 1) return fail(PT_E_INVALID, "%s: np must be >= 1", fn);
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "%s: precision %d", fn, precision);
    if (n < 0 || first_path < 0 || stride < 1) return fail(PT_E_INVALID, "%s: bad path range", fn);
    if (n == 0) return PT_OK;
    const int64_t total = (int64_t)M->nb * M->nk * (nangle - 1) * np;
    if (first_path + (n - 1) * stride >= total)
        return fail(PT_E_INVALID, "%s: path range beyond the last cell", fn);
    PT_HIP(hipSetDevice(c->device));
    std::vector<double> ct, st;
    angle_tables(nangle, vangle, ct, st);
    std::vector<double> ang(ct);
    ang.insert(ang.end(), st.begin(), st.end());
    double *d_ang = nullptr;
    PT_HIP(hipMalloc((void **)&d_ang, 2 * nangle * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d_ang, ang.data(), 2 * nangle * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { (void)hipFree(d_ang); return fail(PT_E_HIP, "pt_replay: %s", hipGetErrorString(e)); }
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    P.cos_t = d_ang;
    P.sin_t = d_ang + nangle;
    P.mfp = mfp;
    P.np = np;
    P.nangle = nangle;
    rc = replay_common(c, P, TRACE_GRID, first_path, n, records, stride, points, max_points, npoints);
    (void)hipFree(d_ang);
    return rc;
}

int pt_replay(pt_ctx *c, const pt_geom *G, const pt_modes *M, double mfp, int32_t nangle,
              const double *vangle, int64_t np, int64_t first_path, int64_t n,
              int32_t precision, pt_path_record *records)
{
    if (n > 0 && !records) return fail(PT_E_INVALID, "pt_replay: records is NULL");
    return grid_replay(c, G, M, mfp, nangle, vangle, np, first_path, 1, n, precision, records,
                       nullptr, 0, nullptr, "pt_replay");
}

int pt_trace_log(pt_ctx *c, const pt_geom *G, const pt_modes *M, double mfp, int32_t nangle,
                 const double *vangle, int64_t np, int64_t first_path, int64_t stride, int64_t n,
                 int32_t precision, int64_t max_points, double *points, int64_t *npoints,
                 pt_path_record *records)
{
    if (n > 0 && (!points || !npoints)) return fail(PT_E_INVALID, "pt_trace_log: points or npoints is NULL");
    if (max_points < 1) return fail(PT_E_INVALID, "pt_trace_log: max_points must be >= 1");
    if (n > 0 && (double)n * (double)max_points * 24.0 > 4.0 * 1073741824.0)
        return fail(PT_E_INVALID, "pt_trace_log: n * max_points * 24 B exceeds 4 GiB");
    return grid_replay(c, G, M, mfp, nangle, vangle, np, first_path, stride, n, precision, records,
                       points, max_points, npoints, "pt_trace_log");
}

int pt_trace_alias(pt_ctx *c, const pt_geom *G, const pt_modes *M, int64_t npaths,
                   int32_t precision, const double *branch_scale, int64_t *mode_paths,
                   int64_t *mode_trans, pt_stats *stats)
{
    if (!c || !G || !M) return fail(PT_E_INVALID, "NULL context, geometry or modes");
    if (G->ctx != c || M->ctx != c) return fail(PT_E_INVALID, "handles belong to another context");
    if (!M->has_alias) return fail(PT_E_INVALID, "modes were created without vel, mfp and weight");
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    if (npaths < 0) return fail(PT_E_INVALID, "npaths must be >= 0");
    if (!mode_paths || !mode_trans) return fail(PT_E_INVALID, "tally arrays are NULL");
    if (stats) memset(stats, 0, sizeof *stats);
    if (npaths == 0) return PT_OK;
    const size_t nm = (size_t)M->nb * M->nk;
    PT_HIP(hipSetDevice(c->device));
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    int rc = set_scale(c, M, branch_scale, P);
    if (rc) return rc;
    P.total = npaths;
    char *wk = nullptr;
    rc = ctx_work(c, align256(2 * nm * sizeof(unsigned long long)), &wk);
    if (rc) return rc;
    unsigned long long *d_tal = (unsigned long long *)wk;
    P.mode_paths = d_tal;
    P.mode_trans = d_tal + nm;
    std::vector<unsigned long long> h_tal(2 * nm);
    unsigned long long host_ctr[kCounterWords] = {0};
    float ms = 0.f;
    int blocks = 0;
    hipError_t e;
    do {
        if ((e = launch_blocks(c, TRACE_ALIAS, precision, 0, &blocks))) break;
        P.chunk = wave_chunk(c, P.total, blocks);
        if ((rc = tail_prepare(c, TRACE_ALIAS, precision, npaths, 0, P))) return rc;
        if ((e = hipMemsetAsync(c->counters, 0, kCounterWords * sizeof(unsigned long long), c->stream))) break;
        if ((e = hipMemsetAsync(d_tal, 0, 2 * nm * sizeof(unsigned long long), c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        LiveTail lt;
        if ((e = launch_trace(P, TRACE_ALIAS, blocks, c->threads, c->count_ops, 0, c->stream, live_tail(c, &lt)))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = hipMemcpyAsync(h_tal.data(), d_tal, 2 * nm * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(host_ctr, c->counters, kCounterWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipStreamSynchronize(c->stream))) break;
        e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_trace_alias: %s", hipGetErrorString(e));
    for (size_t i = 0; i < nm; ++i) {
        mode_paths[i] += (int64_t)h_tal[i];
        mode_trans[i] += (int64_t)h_tal[nm + i];
    }
    fill_stats(stats, host_ctr, ms, 0.f);
    if ((int64_t)host_ctr[1] != npaths)
        return fail(PT_E_HIP, "pt_trace_alias: %lld of %lld paths completed",
                    (long long)host_ctr[1], (long long)npaths);
    return PT_OK;
}

int pt_trace_alias_noalloc(pt_ctx *c, const pt_geom *G, const pt_modes *M, int64_t npaths,
                           int32_t precision, const double *branch_scale, int64_t *mode_paths,
                           int64_t *mode_trans, pt_stats *stats)
{
    if (!c || !G || !M) return fail(PT_E_INVALID, "NULL context, geometry or modes");
    if (G->ctx != c || M->ctx != c) return fail(PT_E_INVALID, "handles belong to another context");
    if (!M->has_alias) return fail(PT_E_INVALID, "modes were created without vel, mfp and weight");
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    if (npaths < 0) return fail(PT_E_INVALID, "npaths must be >= 0");
    if (!mode_paths || !mode_trans) return fail(PT_E_INVALID, "tally arrays are NULL");
    const size_t nm = (size_t)M->nb * M->nk;
    if (nm > kNoAllocMaxModes || M->nb > kNoAllocMaxBranches)
        return fail(PT_E_UNSUPPORTED, "pt_trace_alias_noalloc: %d x %d modes exceed its scratch "
                                      "(at most %zu modes and %d branches)",
                    M->nb, M->nk, kNoAllocMaxModes, kNoAllocMaxBranches);
    if (branch_scale)
        for (int b = 0; b < M->nb; ++b)
            if (!(branch_scale[b] > 0) || !std::isfinite(branch_scale[b]))
                return fail(PT_E_INVALID, "branch_scale[%d] must be positive and finite", b);
    if (!c->noalloc) return fail(PT_E_HIP, "pt_trace_alias_noalloc: no scratch on device %d", c->device);
    if (stats) memset(stats, 0, sizeof *stats);
    if (npaths == 0) return PT_OK;
    PT_HIP(hipSetDevice(c->device));
    // scratch: counters (queue, statistics) | mode_paths | mode_trans | branch scales | tail
    unsigned long long *w = c->noalloc;
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    P.total = npaths;
    P.queue = w;
    P.stats = w + 1;
    P.mode_paths = w + kNoAllocCounters;
    P.mode_trans = P.mode_paths + nm;
    double *d_scale = (double *)(w + kNoAllocCounters + 2 * kNoAllocMaxModes);
    P.m.scale = branch_scale ? d_scale : nullptr;
    // long walks go to the tail records at the end of the scratch
    P.tail_events = c->tail_events;
    if (tail_mode(TRACE_ALIAS, precision) && !c->count_ops && c->tail_max > 0) {
        P.tail = (TailRec *)(w + kNoAllocTailWord);
        P.tail_cap = std::min<int64_t>(c->tail_max, kNoAllocTailRecs);
        P.tail_cap_p = P.tail_cap / 4;
        P.tail_prio_mfp = kTailPrioMfpFrac * std::min(P.g.box[0], std::min(P.g.box[1], P.g.box[2]));
    }
    unsigned long long host_ctr[kCounterWords] = {0};
    float ms = 0.f;
    int blocks = 0;
    hipError_t e;
    do {
        if ((e = launch_blocks(c, TRACE_ALIAS, precision, 0, &blocks))) break;
        P.chunk = wave_chunk(c, P.total, blocks);
        // the counters start at zero (a fill kernel, no memset node); the tallies start at
        // the caller's values, to which the kernel adds
        if ((e = launch_fill(w, kNoAllocCounters, c->stream))) break;
        if ((e = hipMemcpyAsync(P.mode_paths, mode_paths, nm * sizeof(int64_t), hipMemcpyHostToDevice, c->stream))) break;
        if ((e = hipMemcpyAsync(P.mode_trans, mode_trans, nm * sizeof(int64_t), hipMemcpyHostToDevice, c->stream))) break;
        if (branch_scale && (e = hipMemcpyAsync(d_scale, branch_scale, M->nb * sizeof(double), hipMemcpyHostToDevice, c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        if ((e = launch_trace(P, TRACE_ALIAS, blocks, c->threads, c->count_ops, 0, c->stream))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = hipMemcpyAsync(mode_paths, P.mode_paths, nm * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(mode_trans, P.mode_trans, nm * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipMemcpyAsync(host_ctr, w, kNoAllocCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream))) break;
        if ((e = hipStreamSynchronize(c->stream))) break;
        e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_trace_alias_noalloc: %s", hipGetErrorString(e));
    fill_stats(stats, host_ctr, ms, 0.f);
    if ((int64_t)host_ctr[1] != npaths)
        return fail(PT_E_HIP, "pt_trace_alias_noalloc: %lld of %lld paths completed",
                    (long long)host_ctr[1], (long long)npaths);
    return PT_OK;
}

int pt_replay_alias(pt_ctx *c, const pt_geom *G, const pt_modes *M, const double *branch_scale,
                    int64_t first, int64_t n, int32_t precision, pt_path_record *records)
{
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    if (!c || !G || !M) return fail(PT_E_INVALID, "NULL context, geometry or modes");
    if (!M->has_alias) return fail(PT_E_INVALID, "modes were created without vel, mfp and weight");
    if (n < 0 || first < 0) return fail(PT_E_INVALID, "bad path range");
    if (n == 0) return PT_OK;
    if (!records) return fail(PT_E_INVALID, "records is NULL");
    PT_HIP(hipSetDevice(c->device));
    TraceParams P = base_params(c, G, M);
    P.precision = precision;
    int rc = set_scale(c, M, branch_scale, P);
    if (rc) return rc;
    return replay_common(c, P, TRACE_ALIAS, first, n, records);
}

int pt_adjoint_alias(pt_ctx *c, const pt_geom *G, const pt_modes *M, int64_t npaths,
                     int32_t precision, const double *branch_scale, const int32_t *surface_group,
                     int32_t ngroup, double baseline, int32_t nbatch, int64_t *mode_paths,
                     int64_t *mode_trans, double *batch_sums, pt_stats *stats)
{
    if (!batch_sums) return fail(PT_E_INVALID, "batch_sums is NULL");
    return adjoint_common(c, G, M, npaths, precision, branch_scale, surface_group, ngroup,
                          baseline, nbatch, mode_paths, mode_trans, batch_sums, nullptr, stats,
                          "pt_adjoint_alias");
}

int pt_trace_alias_grains(pt_ctx *c, const pt_geom *G, const pt_modes *M, int64_t npaths,
                          int32_t precision, const double *branch_scale, int64_t *mode_paths,
                          int64_t *mode_trans, double *grain_tally, pt_stats *stats)
{
    if (!grain_tally) return fail(PT_E_INVALID, "grain_tally is NULL");
    if (G && G->dg.ngrain_total > 4096) return fail(PT_E_UNSUPPORTED, "more than 4096 grains");
    return adjoint_common(c, G, M, npaths, precision, branch_scale, nullptr, 0, 0.0, 1,
                          mode_paths, mode_trans, nullptr, grain_tally, stats,
                          "pt_trace_alias_grains");
}

}  // extern "C"

// =====================================================================================
// Batches of mesh structures (an extension): one merged geometry, one persistent launch.
// =====================================================================================

struct pt_batch {
    pt_ctx *ctx = nullptr;
    pt_geom *geom = nullptr;               // the merged geometry
    int32_t nstruct = 0, grid = 0, leaf = 0, depth = 0;
    int32_t top_cap = 0;                   // most grains of one structure
    int64_t npaths = 0;
    std::vector<int64_t> path_off;
    std::vector<int32_t> struct_surf_begin; // merged surface range of each structure
    std::vector<void *> owned;
    DevBatch bt{};
    int32_t *d_group = nullptr;            // surface -> 0 (structure facets) or -1 (box)
    // scratch zeroed by every launch: counters(kCounterWords) | sums(kBatchSums n) | finish(n) | start(1) |
    // done(ceil(n/2)) | mode tallies (2 nm)
    unsigned long long *scratch = nullptr;
    size_t scratch_words = 0, nm = 0;
    int64_t nodes = 0, prims = 0;
    // graph of (fill, trace) for one (modes, precision)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t trace_node = nullptr;
    hipGraphNode_t tail_node = nullptr;        // the tail kernel after the trace (or none)
    const TailRec *graph_tail = nullptr;       // tail buffer the graph was built with
    const pt_modes *graph_modes = nullptr;
    int graph_prec = -1;
    int graph_blocks = 0;                  // launch shape the graph was built with
    size_t graph_lds = 0;
    TraceParams graph_params{};
    hipKernelNodeParams graph_knp{};
    void *graph_args[1] = {nullptr};
};

namespace {

struct BvhPrim {
    double lo[3], hi[3], c[3];
    double n[3];        // polygon normal; curved prims face every way
    bool curved;
    int idx;
};

float round_down(double v) { float f = (float)v; if ((double)f > v) f = nextafterf(f, -3.0e38f); return f; }
float round_up(double v) { float f = (float)v; if ((double)f < v) f = nextafterf(f, 3.0e38f); return f; }

// A child slot of a BvhNode: its reference, padded bounds and normal bounds.
struct BvhChild {
    int32_t ref, count;
    float lo[3], hi[3];
    float nlo[3], nhi[3];
};

// Median-split BVH over prims[b, e); returns the child slot of the subtree (a leaf or an
// internal node).  Leaves hold at most `leaf` prims unless the depth limit forces a
// larger leaf.
BvhChild bvh_build(std::vector<BvhPrim> &P, int b, int e, int depth, int max_depth, int leaf,
                   std::vector<BvhNode> &nodes, std::vector<int32_t> &prims, int *depth_out)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], P[i].lo[k]);
            hi[k] = std::max(hi[k], P[i].hi[k]);
            clo[k] = std::min(clo[k], P[i].c[k]);
            chi[k] = std::max(chi[k], P[i].c[k]);
        }
    BvhChild me{};
    for (int k = 0; k < 3; ++k) {
        // outward padding: 1e-3 nm + 1e-6 relative covers FP32 rounding of positions
        const float l = round_down(lo[k]), h = round_up(hi[k]);
        me.lo[k] = l - (1.0e-3f + 1.0e-6f * fabsf(l));
        me.hi[k] = h + (1.0e-3f + 1.0e-6f * fabsf(h));
        double nl = 1e300, nh = -1e300;
        for (int i = b; i < e; ++i) {
            nl = std::min(nl, P[i].curved ? -1.0e30 : P[i].n[k]);
            nh = std::max(nh, P[i].curved ? 1.0e30 : P[i].n[k]);
        }
        me.nlo[k] = round_down(std::max(nl, -1.0e30));
        me.nhi[k] = round_up(std::min(nh, 1.0e30));
    }
    *depth_out = std::max(*depth_out, depth);
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    if (e - b <= leaf || depth >= max_depth || !(chi[axis] > clo[axis])) {
        me.ref = -(int32_t)prims.size() - 1;
        me.count = e - b;
        for (int i = b; i < e; ++i) prims.push_back(P[i].idx);
        return me;
    }
    const int ni = (int)nodes.size();
    nodes.push_back(BvhNode{});
    const int m = (b + e) / 2;
    std::nth_element(P.begin() + b, P.begin() + m, P.begin() + e,
                     [axis](const BvhPrim &x, const BvhPrim &y) { return x.c[axis] < y.c[axis]; });
    const BvhChild kid[2] = {bvh_build(P, b, m, depth + 1, max_depth, leaf, nodes, prims, depth_out),
                             bvh_build(P, m, e, depth + 1, max_depth, leaf, nodes, prims, depth_out)};
    for (int c = 0; c < 2; ++c) {
        nodes[ni].child[c] = kid[c].ref;
        nodes[ni].count[c] = kid[c].count;
        for (int k = 0; k < 3; ++k) {
            nodes[ni].lo[c][k] = kid[c].lo[k];
            nodes[ni].hi[c][k] = kid[c].hi[k];
            nodes[ni].nlo[c][k] = kid[c].nlo[k];
            nodes[ni].nhi[c][k] = kid[c].nhi[k];
        }
    }
    me.ref = ni;
    me.count = 0;
    return me;
}

// Root node of a BVH: the top split, or for a single leaf a node whose second child is
// an empty leaf (no prims, inverted bounds).
int bvh_root(std::vector<BvhPrim> &P, int max_depth, int leaf, std::vector<BvhNode> &nodes,
             std::vector<int32_t> &prims, int *depth_out)
{
    const BvhChild top = bvh_build(P, 0, (int)P.size(), 0, max_depth, leaf, nodes, prims, depth_out);
    if (top.ref >= 0) return top.ref;
    BvhNode nd{};
    nd.child[0] = top.ref;
    nd.count[0] = top.count;
    nd.child[1] = -1;
    nd.count[1] = 0;
    for (int k = 0; k < 3; ++k) {
        nd.lo[0][k] = top.lo[k];
        nd.hi[0][k] = top.hi[k];
        nd.nlo[0][k] = top.nlo[k];
        nd.nhi[0][k] = top.nhi[k];
        nd.lo[1][k] = 3.0e38f;
        nd.hi[1][k] = -3.0e38f;
        nd.nlo[1][k] = -1.0e30f;
        nd.nhi[1][k] = 1.0e30f;
    }
    nodes.push_back(nd);
    return (int)nodes.size() - 1;
}

void prim_bounds(const pt_geom_desc &d, const pt_surface &s, BvhPrim &p)
{
    p.curved = s.kind != PT_SHAPE_POLYGON;
    for (int k = 0; k < 3; ++k) p.n[k] = s.normal[k];
    if (s.kind == PT_SHAPE_POLYGON) {
        for (int k = 0; k < 3; ++k) { p.lo[k] = 1e300; p.hi[k] = -1e300; }
        for (int v = 0; v < s.nvertex; ++v)
            for (int k = 0; k < 3; ++k) {
                const double x = d.vertices[3 * (s.vertex_offset + v) + k];
                p.lo[k] = std::min(p.lo[k], x);
                p.hi[k] = std::max(p.hi[k], x);
            }
    } else {
        for (int k = 0; k < 3; ++k) {
            p.lo[k] = s.center[k] - s.radius;
            p.hi[k] = s.center[k] + s.radius;
        }
        if (s.kind == PT_SHAPE_CYLINDER) { p.lo[2] = 0.0; p.hi[2] = d.box[2]; }
    }
    for (int k = 0; k < 3; ++k) p.c[k] = 0.5 * (p.lo[k] + p.hi[k]);
}

bool same_walls(const pt_geom_desc &a, const pt_geom_desc &b)
{
    for (int k = 0; k < 3; ++k)
        if (a.box[k] != b.box[k]) return false;
    if (a.scope[0] != b.scope[0] || a.scope[1] != b.scope[1] || a.nmaterial != b.nmaterial)
        return false;
    for (int w = 0; w < 6; ++w) {
        const pt_surface &x = a.surfaces[w], &y = b.surfaces[w];
        if (x.specularity != y.specularity || x.trans_kind != y.trans_kind ||
            x.trans_param != y.trans_param)
            return false;
    }
    return true;
}

}  // namespace

extern "C" {

int pt_batch_create(pt_ctx *c, int32_t nstruct, const pt_geom_desc *const *descs,
                    const int64_t *npaths, int32_t leaf_size, int32_t grid, pt_batch **out)
{
    if (!c || !descs || !npaths || !out) return fail(PT_E_INVALID, "pt_batch_create: NULL argument");
    *out = nullptr;
    if (nstruct < 1 || nstruct > 4096) return fail(PT_E_INVALID, "nstruct must be in 1..4096");
    if (leaf_size < 1 || leaf_size > 64) return fail(PT_E_INVALID, "leaf_size must be in 1..64");
    if (grid < 1 || grid > 1024) return fail(PT_E_INVALID, "grid must be in 1..1024");
    if (c->threads != 256) return fail(PT_E_UNSUPPORTED, "batch traces run 256-lane workgroups");
    const pt_geom_desc &d0 = *descs[0];
    int64_t ntot_paths = 0;
    size_t ns = 6, nv = 0;
    int ng = 0;
    for (int s = 0; s < nstruct; ++s) {
        const pt_geom_desc &d = *descs[s];
        if (d.kind_grain != PT_GRAIN_MESH) return fail(PT_E_UNSUPPORTED, "structure %d is not a mesh", s);
        if (d.ngrain_total != d.ngrain + 1 || d.grain_offset[1] != 6)
            return fail(PT_E_GEOMETRY, "structure %d: grain 0 must be the six box walls", s);
        if (!same_walls(d0, d))
            return fail(PT_E_GEOMETRY, "structure %d: box, scope, materials and walls must be the "
                                       "batch's", s);
        if (npaths[s] < 0 || ntot_paths > ((int64_t)1 << 53) - npaths[s])
            return fail(PT_E_INVALID, "bad path count of structure %d", s);
        ntot_paths += npaths[s];
        ns += (size_t)(d.nsurface - 6);
        nv += (size_t)d.nvertex;
        ng += d.ngrain;
    }
    // merged descriptor: shared box grain, then every structure's grains
    std::vector<pt_surface> surf(ns);
    std::vector<double> vtx;
    vtx.reserve(3 * nv + 3 * 4);
    std::vector<int32_t> goff(1, 0), ginc(1, d0.grain_incident[0]);
    std::vector<double> gen(d0.generatrix, d0.generatrix + 3);
    for (int w = 0; w < 6; ++w) {
        surf[w] = d0.surfaces[w];
        if (surf[w].kind == PT_SHAPE_POLYGON) {
            const int64_t vo = (int64_t)vtx.size() / 3;
            for (int i = 0; i < 3 * surf[w].nvertex; ++i)
                vtx.push_back(d0.vertices[3 * surf[w].vertex_offset + i]);
            surf[w].vertex_offset = vo;
        }
    }
    goff.push_back(6);
    pt_batch *B = new pt_batch();
    B->ctx = c;
    B->nstruct = nstruct;
    B->grid = grid;
    B->leaf = leaf_size;
    B->npaths = ntot_paths;
    B->path_off.assign(nstruct + 1, 0);
    B->struct_surf_begin.assign(nstruct + 1, 6);
    size_t k = 6;
    int gbase = 1;
    for (int s = 0; s < nstruct; ++s) {
        const pt_geom_desc &d = *descs[s];
        B->path_off[s + 1] = B->path_off[s] + npaths[s];
        B->struct_surf_begin[s] = (int32_t)k;
        const int64_t vbase = (int64_t)vtx.size() / 3;
        vtx.insert(vtx.end(), d.vertices, d.vertices + 3 * d.nvertex);
        auto remap = [gbase](int x) { return x >= 1 ? gbase + x - 1 : x; };
        for (int g = 1; g <= d.ngrain; ++g) {
            for (int i = d.grain_offset[g]; i < d.grain_offset[g + 1]; ++i) {
                pt_surface t = d.surfaces[i];
                t.grain_inside = remap(t.grain_inside);
                t.grain_outside = remap(t.grain_outside);
                if (t.kind == PT_SHAPE_POLYGON) t.vertex_offset += vbase;
                surf[k++] = t;
            }
            goff.push_back((int32_t)k);
            ginc.push_back(d.grain_incident[g]);
            gen.insert(gen.end(), d.generatrix + 3 * g, d.generatrix + 3 * g + 3);
        }
        gbase += d.ngrain;
    }
    B->struct_surf_begin[nstruct] = (int32_t)k;
    pt_geom_desc M = d0;
    M.ngrain = ng;
    M.ngrain_total = ng + 1;
    M.igrain_fill = -1;
    M.nsurface = (int32_t)ns;
    M.nvertex = (int64_t)vtx.size() / 3;
    M.grain_offset = goff.data();
    M.grain_incident = ginc.data();
    M.generatrix = gen.data();
    M.surfaces = surf.data();
    M.vertices = vtx.data();
    int rc = pt_geom_create(c, &M, &B->geom);
    if (rc) { delete B; return rc; }

    // BVH per grain over its candidate (internal) surfaces
    std::vector<BvhNode> nodes;
    std::vector<int32_t> prims, root(ng + 1, -1);
    for (int g = 1; g <= ng; ++g) {
        std::vector<BvhPrim> P;
        for (int i = goff[g]; i < goff[g + 1]; ++i) {
            if (!surf[i].internal) continue;
            BvhPrim p;
            prim_bounds(M, surf[i], p);
            p.idx = i;
            P.push_back(p);
        }
        if (P.empty()) continue;
        int dep = 0;
        root[g] = bvh_root(P, kBvhStack - 2, leaf_size, nodes, prims, &dep);
        B->depth = std::max(B->depth, dep);
    }
    B->nodes = (int64_t)nodes.size();
    B->prims = (int64_t)prims.size();
    if (B->nodes >= kLdsRef) {
        pt_batch_destroy(B);
        return fail(PT_E_UNSUPPORTED, "pt_batch_create: %lld BVH nodes (at most %d)",
                    (long long)nodes.size(), kLdsRef - 1);
    }
    // image of the top BVH levels for LDS staging: per grain the first kTopImage heap
    // slots (root, its children, theirs) as stored (global child references); a grain
    // without candidates gets a root with two empty leaves
    std::vector<BvhNode> top((size_t)(ng + 1) * kTopImage, BvhNode{});
    for (int g = 1; g <= ng; ++g) {
        BvhNode *h = &top[(size_t)g * kTopImage];
        if (root[g] < 0) {
            for (int c = 0; c < 2; ++c) {
                h[0].child[c] = -1;
                h[0].count[c] = 0;
                for (int k = 0; k < 3; ++k) {
                    h[0].lo[c][k] = 3.0e38f;
                    h[0].hi[c][k] = -3.0e38f;
                    h[0].nlo[c][k] = -1.0e30f;
                    h[0].nhi[c][k] = 1.0e30f;
                }
            }
            continue;
        }
        std::vector<int32_t> ref(kTopImage, -1);
        ref[0] = root[g];
        for (int j = 0; j < kTopImage; ++j) {
            if (ref[j] < 0) continue;
            h[j] = nodes[(size_t)ref[j]];
            for (int c = 0; c < 2; ++c)
                if (2 * j + 1 + c < kTopImage && h[j].child[c] >= 0) ref[2 * j + 1 + c] = h[j].child[c];
        }
    }
    std::vector<int32_t> sgrain(nstruct + 1, 1);
    int max_ng = 0;
    for (int s = 0; s < nstruct; ++s) {
        sgrain[s + 1] = sgrain[s] + descs[s]->ngrain;
        max_ng = std::max(max_ng, descs[s]->ngrain);
    }
    B->top_cap = max_ng;
    // start locator: per structure, incident polygons overlapping each (y, z) cell
    std::vector<int32_t> loc_off(1, 0), loc_idx;
    for (int s = 0; s < nstruct; ++s) {
        std::vector<std::vector<int32_t>> cells((size_t)grid * grid);
        for (int i = B->struct_surf_begin[s]; i < B->struct_surf_begin[s + 1]; ++i) {
            const pt_surface &t = surf[i];
            if (!t.incident || t.kind != PT_SHAPE_POLYGON) continue;
            BvhPrim p;
            prim_bounds(M, t, p);
            int y0 = (int)std::floor(p.lo[1] / M.box[1] * grid) - 1, y1 = (int)std::floor(p.hi[1] / M.box[1] * grid) + 1;
            int z0 = (int)std::floor(p.lo[2] / M.box[2] * grid) - 1, z1 = (int)std::floor(p.hi[2] / M.box[2] * grid) + 1;
            y0 = std::max(y0, 0); z0 = std::max(z0, 0);
            y1 = std::min(y1, grid - 1); z1 = std::min(z1, grid - 1);
            for (int y = y0; y <= y1; ++y)
                for (int z = z0; z <= z1; ++z) cells[(size_t)y * grid + z].push_back(i);
        }
        for (auto &cl : cells) {
            loc_idx.insert(loc_idx.end(), cl.begin(), cl.end());
            loc_off.push_back((int32_t)loc_idx.size());
        }
    }
    // per-XCD queues: contiguous structure ranges of about equal path counts
    std::vector<int64_t> qoff(kXcdQueues + 1, 0);
    qoff[kXcdQueues] = ntot_paths;
    for (int q = 1, s = 0; q < kXcdQueues; ++q) {
        const int64_t target = (int64_t)((double)ntot_paths * q / kXcdQueues);
        while (s < nstruct && B->path_off[s] < target) ++s;
        qoff[q] = std::max(qoff[q - 1], B->path_off[s]);
    }
    std::vector<int32_t> group(ns, -1);
    for (size_t i = 6; i < ns; ++i) group[i] = 0;
    B->nm = 0;
    const size_t n = (size_t)nstruct;
    B->scratch_words = kCounterWords + kBatchSums * n + n + 1 + (n + 1) / 2;
    hipStream_t st = c->stream;
    int64_t *d_poff = nullptr, *d_qoff = nullptr;
    int32_t *d_root = nullptr, *d_prims = nullptr, *d_loff = nullptr, *d_lidx = nullptr;
    int32_t *d_sgrain = nullptr;
    BvhNode *d_nodes = nullptr, *d_top = nullptr;
    double *d_base = nullptr;
    std::vector<double> zeros(n, 0.0);
    if ((rc = upload(B->path_off, &d_poff, st)) || (rc = upload(root, &d_root, st)) ||
        (rc = upload(nodes, &d_nodes, st)) || (rc = upload(prims, &d_prims, st)) ||
        (rc = upload(loc_off, &d_loff, st)) || (rc = upload(loc_idx, &d_lidx, st)) ||
        (rc = upload(group, &B->d_group, st)) || (rc = upload(zeros, &d_base, st)) ||
        (rc = upload(qoff, &d_qoff, st)) || (rc = upload(top, &d_top, st)) ||
        (rc = upload(sgrain, &d_sgrain, st))) {
        B->owned = {d_poff, d_root, d_nodes, d_prims, d_loff, d_lidx, B->d_group, d_base, d_qoff,
                    d_top, d_sgrain};
        pt_batch_destroy(B);
        return rc;
    }
    B->owned = {d_poff, d_root, d_nodes, d_prims, d_loff, d_lidx, B->d_group, d_base, d_qoff,
                d_top, d_sgrain};
    DevBatch &bt = B->bt;
    bt.nstruct = nstruct;
    bt.grid = grid;
    bt.path_off = d_poff;
    bt.grain_root = d_root;
    bt.nodes = d_nodes;
    bt.prims = d_prims;
    bt.loc_off = d_loff;
    bt.loc_idx = d_lidx;
    bt.baseline = d_base;
    bt.qoff = d_qoff;
    bt.top = d_top;
    bt.sgrain = d_sgrain;
    bt.top_k = 0;
    bt.top_cap = B->top_cap;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        pt_batch_destroy(B);
        return fail(PT_E_HIP, "pt_batch_create: %s", hipGetErrorString(e));
    }
    *out = B;
    return PT_OK;
}

int pt_batch_destroy(pt_batch *B)
{
    if (!B) return PT_OK;
    (void)hipSetDevice(B->ctx->device);
    (void)hipStreamSynchronize(B->ctx->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    for (void *p : B->owned)
        if (p) (void)hipFree(p);
    if (B->scratch) (void)hipFree(B->scratch);
    if (B->geom) pt_geom_destroy(B->geom);
    delete B;
    return PT_OK;
}

int pt_batch_info(const pt_batch *B, int64_t *info)
{
    if (!B || !info) return fail(PT_E_INVALID, "pt_batch_info: NULL argument");
    info[0] = B->geom->dg.nsurface;
    info[1] = B->geom->dg.ngrain_total;
    info[2] = B->nodes;
    info[3] = B->prims;
    info[4] = B->depth;
    info[5] = B->npaths;
    return PT_OK;
}

int pt_batch_set_specularity(pt_batch *B, const double *p)
{
    if (!B || !p) return fail(PT_E_INVALID, "pt_batch_set_specularity: NULL argument");
    std::vector<double> all(B->geom->dg.nsurface);
    for (int i = 0; i < 6; ++i) all[i] = B->geom->surf[i].p;
    for (int s = 0; s < B->nstruct; ++s) {
        if (!(p[s] > 0.0 && p[s] < 1.0))
            return fail(PT_E_INVALID, "mean specularity of structure %d is %g: need 0 < p < 1", s, p[s]);
        for (int i = B->struct_surf_begin[s]; i < B->struct_surf_begin[s + 1]; ++i) all[i] = p[s];
    }
    return pt_geom_set_specularity(B->geom, all.data());
}

int pt_batch_set_top_levels(pt_batch *B, int32_t levels)
{
    if (!B) return fail(PT_E_INVALID, "pt_batch_set_top_levels: NULL batch");
    if (levels < 0 || levels > kMaxTopLevels)
        return fail(PT_E_INVALID, "pt_batch_set_top_levels: levels must be in 0..%d", kMaxTopLevels);
    B->bt.top_k = (1 << levels) - 1;
    return PT_OK;
}

}  // extern "C"

namespace {

int batch_prepare(pt_batch *B, const pt_modes *M, int precision, TraceParams &P, size_t *lds,
                  int *blocks)
{
    pt_ctx *c = B->ctx;
    if (!M || M->ctx != c) return fail(PT_E_INVALID, "modes belong to another context");
    if (!M->has_alias) return fail(PT_E_INVALID, "modes were created without vel, mfp and weight");
    if (precision != PT_PREC_FP64 && precision != PT_PREC_FP32_RETEST &&
        precision != PT_PREC_FP32_FILTER)
        return fail(PT_E_UNSUPPORTED, "precision %d", precision);
    const size_t nm = (size_t)M->nb * M->nk;
    const size_t S = (size_t)kBatchSums * B->nstruct;
    const size_t need = kCounterWords + S + B->nstruct + 1 + (B->nstruct + 1) / 2 + 2 * nm + kXcdQueues;
    if (B->scratch_words < need || !B->scratch) {
        if (B->scratch) PT_HIP(hipFree(B->scratch));
        B->scratch = nullptr;
        PT_HIP(hipMalloc((void **)&B->scratch, need * sizeof(unsigned long long)));
        B->scratch_words = need;
        if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
        if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
    }
    B->nm = nm;
    unsigned long long *w = B->scratch;
    P = base_params(c, B->geom, M);
    P.precision = precision;
    P.queue = w;
    P.stats = w + 1;
    P.total = B->npaths;
    P.ngroup = 1;
    P.nbatch = 1;
    P.surf_group = B->d_group;
    P.baseline = 0.0;
    P.m.scale = nullptr;
    DevBatch bt = B->bt;
    bt.sums = (double *)(w + kCounterWords);
    bt.finish = w + kCounterWords + S;
    bt.start = w + kCounterWords + S + B->nstruct;
    bt.done = (unsigned int *)(w + kCounterWords + S + B->nstruct + 1);
    P.bt = bt;
    P.mode_paths = w + kCounterWords + S + B->nstruct + 1 + (B->nstruct + 1) / 2;
    P.mode_trans = P.mode_paths + nm;
    P.bt.qctr = P.mode_trans + nm;
    P.adj = nullptr;
    *lds = adj_lds_bytes(P, c->threads);
    hipDeviceProp_t prop;
    PT_HIP(hipGetDeviceProperties(&prop, c->device));
    if (*lds + trace_static_lds(TRACE_BATCH) > prop.sharedMemPerBlock)
        return fail(PT_E_UNSUPPORTED, "batch of %d structures needs %zu bytes of LDS per workgroup "
                    "(fewer top BVH levels in LDS)", B->nstruct, *lds + trace_static_lds(TRACE_BATCH));
    PT_HIP(launch_blocks(c, TRACE_BATCH, precision, *lds, blocks));
    return tail_prepare(c, TRACE_BATCH, precision, B->npaths, 1, P);
}

int batch_finish(pt_batch *B, double *sums, double *finish_ms, pt_stats *stats, float ms)
{
    pt_ctx *c = B->ctx;
    const size_t n = (size_t)B->nstruct;
    const size_t S = (size_t)kBatchSums * n;
    std::vector<unsigned long long> h(kCounterWords + S + n + 1);
    PT_HIP(hipMemcpyAsync(h.data(), B->scratch, h.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
    PT_HIP(hipStreamSynchronize(c->stream));
    const double *hs = (const double *)(h.data() + kCounterWords);
    if (sums)
        for (size_t i = 0; i < S; ++i) sums[i] += hs[i];
    const unsigned long long t0 = h[kCounterWords + S + n];
    if (finish_ms)
        for (size_t s = 0; s < n; ++s) {
            const unsigned long long t = h[kCounterWords + S + s];
            // wall_clock64 ticks at 100 MHz on gfx9
            finish_ms[s] = t >= t0 && t ? (double)(t - t0) * 1.0e-5 : -1.0;
        }
    fill_stats(stats, h.data(), ms, 0.f);
    if ((int64_t)h[1] != B->npaths)
        return fail(PT_E_HIP, "batch trace: %lld of %lld paths completed", (long long)h[1],
                    (long long)B->npaths);
    return PT_OK;
}

}  // namespace

extern "C" {

int pt_batch_trace(pt_ctx *c, pt_batch *B, const pt_modes *M, int32_t precision,
                   const double *baseline, double *sums, double *finish_ms, pt_stats *stats)
{
    if (!c || !B || B->ctx != c) return fail(PT_E_INVALID, "pt_batch_trace: bad context or batch");
    PT_HIP(hipSetDevice(c->device));
    TraceParams P;
    size_t lds = 0;
    int blocks = 0;
    int rc = batch_prepare(B, M, precision, P, &lds, &blocks);
    if (rc) return rc;
    if (stats) memset(stats, 0, sizeof *stats);
    std::vector<double> base(B->nstruct, 0.0);
    if (baseline) base.assign(baseline, baseline + B->nstruct);
    float ms = 0.f;
    hipError_t e;
    do {
        if ((e = hipMemcpyAsync((void *)B->bt.baseline, base.data(), base.size() * sizeof(double),
                                hipMemcpyHostToDevice, c->stream))) break;
        if ((e = launch_fill(B->scratch, B->scratch_words, c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        LiveTail lt;
        if ((e = launch_trace(P, TRACE_BATCH, blocks, c->threads, c->count_ops, lds, c->stream, live_tail(c, &lt)))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = hipEventSynchronize(c->ev1))) break;
        e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_batch_trace: %s", hipGetErrorString(e));
    return batch_finish(B, sums, finish_ms, stats, ms);
}

// The (zero-fill, trace) pair as a hipGraph: built once per (modes, precision),
// replayed with the context's current seed (the trace node's parameters are updated).
int pt_batch_trace_graph(pt_ctx *c, pt_batch *B, const pt_modes *M, int32_t precision,
                         const double *baseline, double *sums, double *finish_ms, pt_stats *stats)
{
    if (!c || !B || B->ctx != c) return fail(PT_E_INVALID, "pt_batch_trace_graph: bad context or batch");
    PT_HIP(hipSetDevice(c->device));
    TraceParams P;
    size_t lds = 0;
    int blocks = 0;
    int rc = batch_prepare(B, M, precision, P, &lds, &blocks);
    if (rc) return rc;
    // the graph has no pre-pass node: its queue runs the paths in order
    P.first_ids = nullptr;
    if (stats) memset(stats, 0, sizeof *stats);
    if (!B->exec || B->graph_modes != M || B->graph_prec != precision ||
        B->graph_blocks != blocks || B->graph_lds != lds || B->graph_tail != P.tail) {
        if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
        if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
        PT_HIP(hipGraphCreate(&B->graph, 0));
        hipKernelNodeParams fk{};
        unsigned long long *fill_ptr = B->scratch;
        size_t fill_n = B->scratch_words;
        void *fargs[2] = {(void *)&fill_ptr, (void *)&fill_n};      // copied by the node
        fk.func = (void *)fill_kernel_fn();
        fk.gridDim = dim3(1024);
        fk.blockDim = dim3(256);
        fk.sharedMemBytes = 0;
        fk.kernelParams = fargs;
        fk.extra = nullptr;
        hipGraphNode_t fnode = nullptr;
        PT_HIP(hipGraphAddKernelNode(&fnode, B->graph, nullptr, 0, &fk));
        B->graph_params = P;
        B->graph_args[0] = (void *)&B->graph_params;
        hipKernelNodeParams &k = B->graph_knp;
        k.func = (void *)trace_kernel_fn(TRACE_BATCH, precision);
        k.gridDim = dim3(blocks);
        k.blockDim = dim3(c->threads);
        k.sharedMemBytes = (unsigned)lds;
        k.kernelParams = B->graph_args;
        k.extra = nullptr;
        PT_HIP(hipGraphAddKernelNode(&B->trace_node, B->graph, &fnode, 1, &k));
        B->tail_node = nullptr;
        if (P.tail_cap > 0) {
            // the long walks the trace handed over, in the same graph
            hipKernelNodeParams t = k;
            t.func = (void *)tail_kernel_fn(TRACE_BATCH, precision);
            t.sharedMemBytes = 0;
            PT_HIP(hipGraphAddKernelNode(&B->tail_node, B->graph, &B->trace_node, 1, &t));
        }
        PT_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->graph_modes = M;
        B->graph_prec = precision;
        B->graph_blocks = blocks;
        B->graph_lds = lds;
        B->graph_tail = P.tail;
    }
    B->graph_params = P;                       // the current seed and tuning
    B->graph_args[0] = (void *)&B->graph_params;
    PT_HIP(hipGraphExecKernelNodeSetParams(B->exec, B->trace_node, &B->graph_knp));
    if (B->tail_node) {
        hipKernelNodeParams t = B->graph_knp;
        t.func = (void *)tail_kernel_fn(TRACE_BATCH, precision);
        t.sharedMemBytes = 0;
        PT_HIP(hipGraphExecKernelNodeSetParams(B->exec, B->tail_node, &t));
    }
    std::vector<double> base(B->nstruct, 0.0);
    if (baseline) base.assign(baseline, baseline + B->nstruct);
    float ms = 0.f;
    hipError_t e;
    do {
        if ((e = hipMemcpyAsync((void *)B->bt.baseline, base.data(), base.size() * sizeof(double),
                                hipMemcpyHostToDevice, c->stream))) break;
        if ((e = hipEventRecord(c->ev0, c->stream))) break;
        if ((e = hipGraphLaunch(B->exec, c->stream))) break;
        if ((e = hipEventRecord(c->ev1, c->stream))) break;
        if ((e = hipEventSynchronize(c->ev1))) break;
        e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    } while (0);
    if (e != hipSuccess) return fail(PT_E_HIP, "pt_batch_trace_graph: %s", hipGetErrorString(e));
    return batch_finish(B, sums, finish_ms, stats, ms);
}

}  // extern "C"
