// pt_trace.hip -- gfx950 kernels of the phonon Monte Carlo trace (FP64 path).
//
// One phonon path per lane.  The event loop restates calcTransmit
// (src/RTMCphon.f90:1608-2414) and the geometry kernels (RT:2883-3301) operation by
// operation in FP64; this file is compiled with -ffp-contract=off so that every lane
// computes the same bits as a sequential CPU run on the same Philox stream.
//
// trace_grid_kernel is persistent: each wave takes chunks of path ids from one global
// counter and refills lanes whose path has ended (ballot + prefix count), so lanes
// with long (diffusive) paths do not hold idle lanes of the same wave hostage.
#include "pt_internal.h"
#include "pt_math.h"
#include "pt_philox.h"

namespace pt {

namespace {

constexpr double kTwoPi = 3.14159265358979323846 * 2.0;    // M_TWOPI, RT:63
constexpr double kSmallVal = 1.0e-5;                         // SMALL_VAL, RT:69
constexpr double kBallistic = 1.0e10;                        // RT:1686
#define PT_NUDGE ((double)1.0e-7f)                           // REAL(4) literal, RT:1779
#define PT_SPHERE_EPS ((double)1e-10f)                       // REAL(4) literal, RT:3001

// Algorithmic FP64 operation count of the trace (roofline accounting).  Every add,
// sub, mul, div and sqrt of the restated arithmetic counts one; comparisons, sign
// flips by selection and integer work do not.  The per-branch constants below are
// counted off the code they sit next to.  The timed kernel never stores its OpCount,
// so the compiler removes the counting; trace_grid_kernel<true> writes it out.
struct OpCount {
    unsigned long long flops;
    unsigned long long philox;     // Philox4x32-10 blocks (integer VALU work)
    unsigned long long nodes;      // BVH nodes visited (scan and ball)
    unsigned long long tests;      // BVH leaf surfaces tested in FP64
    // the lane's event statistics in LDS (counter k at st[256 * k]): FP64 retests of the
    // filter at k = 6, full-scan fallbacks at k = 7
    unsigned long long *st;
};

constexpr unsigned kFlopsLog = 33;     // pt_log: 4 + Horner 20 + 5 + 4
constexpr unsigned kFlopsSincos = 46;  // pt_sincos: reduction 9, sin 18, cos 19

// Draw d of the lane's path: pt_philox_draw(seed, path, d), one Philox block per
// two draws.
struct Rng {
    uint64_t seed;
    uint64_t path;
    uint32_t draw;
    double spare;
    OpCount *oc;
    __device__ void reset(uint64_t p)
    {
        path = p;
        draw = 0;
    }
    __device__ double next()
    {
        double u;
        if (draw & 1u) {
            u = spare;
        } else {
            double b[2];
            oc->philox++;
            pt_philox_block(seed, path, (uint64_t)(draw >> 1), b);
            u = b[0];
            spare = b[1];
        }
        ++draw;
        return u;
    }
};

__device__ inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// calcSurfaceDistancePolygon, RT:2907-2959: one-sided plane hit, then the convex
// point-in-polygon test with on-edge counted inside.
// Geometry read through the constant address space: with a wave-uniform index the
// loads become scalar (SMEM) loads into SGPRs.
typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(4))) DSurf CSurf;
typedef const __attribute__((address_space(4))) int32_t cint;

// DSurf record through a scalar (constant address space) load.
__device__ inline DSurf load_surf_c(const DSurf *base, int idx)
{
    CSurf *p = (CSurf *)base + idx;
    DSurf o;
    for (int k = 0; k < 3; ++k) { o.n[k] = p->n[k]; o.c[k] = p->c[k]; }
    o.r = p->r;
    o.p = p->p;
    o.tparam = p->tparam;
    o.kind = p->kind;
    o.nv = p->nv;
    o.voff = p->voff;
    o.gout = p->gout;
    o.tkind = p->tkind;
    return o;
}


// calcSurfaceDistancePolygon, RT:2907-2959, in two parts: the one-sided plane hit
// (cosang > 0, l >= 0), then the convex point-in-polygon test of the hit point with
// on-edge counted inside.
template <class NT, class VT>
__device__ inline bool plane_hit(NT *n, VT *v, const double *rs, const double *nd, double *lhit,
                                 OpCount &oc)
{
    const double cosang = nd[0] * n[0] + nd[1] * n[1] + nd[2] * n[2];
    oc.flops += 5;
    if (cosang <= 0.0) return false;
    oc.flops += 9;
    double d[3] = {v[0] - rs[0], v[1] - rs[1], v[2] - rs[2]};
    const double l = (d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) / cosang;
    if (l < 0.0) return false;
    *lhit = l;
    return true;
}

template <class VT>
__device__ inline bool polygon_plane(VT *vtx, const DSurf &s, const double *rs, const double *nd,
                                     double *lhit, OpCount &oc)
{
    return plane_hit(s.n, vtx + 3 * (int64_t)s.voff, rs, nd, lhit, oc);
}

template <class VT>
__device__ inline bool polygon_contains_at(VT *v, int nv, const double *rs, const double *nd,
                                           double l, OpCount &oc)
{
    double h[3];
    for (int k = 0; k < 3; ++k) h[k] = rs[k] + nd[k] * l;
    oc.flops += 6;
    double c1[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nv; ++i) {
        VT *vi = v + 3 * i;
        VT *vn = v + 3 * (i + 1 == nv ? 0 : i + 1);
        double ra[3], rb[3], ci[3];
        for (int k = 0; k < 3; ++k) { ra[k] = vn[k] - vi[k]; rb[k] = h[k] - vi[k]; }
        cross3(ra, rb, ci);
        oc.flops += (i == 0) ? 15 : 20;
        if (i == 0) {
            c1[0] = ci[0]; c1[1] = ci[1]; c1[2] = ci[2];
        } else if (dot3(c1, ci) < 0.0) {
            return false;
        }
    }
    return true;
}

template <class VT>
__device__ inline bool polygon_contains(VT *vtx, const DSurf &s, const double *rs,
                                        const double *nd, double l, OpCount &oc)
{
    return polygon_contains_at(vtx + 3 * (int64_t)s.voff, s.nv, rs, nd, l, oc);
}

template <class VT>
__device__ bool dist_polygon_v(VT *vtx, const DSurf &s, const double *rs, const double *nd,
                               double *lhit, OpCount &oc)
{
    double l;
    if (!polygon_plane(vtx, s, rs, nd, &l, oc) || !polygon_contains(vtx, s, rs, nd, l, oc))
        return false;
    *lhit = l;
    return true;
}

__device__ inline bool dist_polygon(const DevGeom &g, const DSurf &s, const double *rs,
                                    const double *nd, double *lhit, OpCount &oc)
{
    return dist_polygon_v(g.vtx, s, rs, nd, lhit, oc);
}

// isInsidePolygon, RT:2961-2983
template <class VT>
__device__ bool inside_polygon(VT *v, int nv, const double *p, OpCount &oc)
{
    double c1[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nv; ++i) {
        VT *vi = v + 3 * i;
        VT *vn = v + 3 * (i + 1 == nv ? 0 : i + 1);
        double ra[3], rb[3], ci[3];
        for (int k = 0; k < 3; ++k) { ra[k] = vn[k] - vi[k]; rb[k] = p[k] - vi[k]; }
        cross3(ra, rb, ci);
        oc.flops += (i == 0) ? 15 : 20;
        if (i == 0) {
            c1[0] = ci[0]; c1[1] = ci[1]; c1[2] = ci[2];
        } else if (dot3(c1, ci) < 0.0) {
            return false;
        }
    }
    return true;
}

// calcSurfaceDistanceSphere, RT:2985-3038
__device__ bool dist_sphere(const DSurf &s, const double *rs, const double *nd, double *lhit,
                            OpCount &oc)
{
    double rsc[3] = {s.c[0] - rs[0], s.c[1] - rs[1], s.c[2] - rs[2]};
    const double rd = dot3(rsc, nd);
    const double rr = dot3(rsc, rsc);
    oc.flops += 15;
    if (rr > s.r * s.r - PT_SPHERE_EPS) {
        if (rd <= 0.0) return false;
    }
    const double d = rd * rd + s.r * s.r - rr;
    oc.flops += 4;
    if (d < 0.0) return false;
    oc.flops += 3;
    const double sq = sqrt(d);
    const double l1 = rd - sq, l2 = rd + sq;
    if (l1 > 0.0) *lhit = l1;
    else if (l2 > 0.0) *lhit = l2;
    else return false;
    return true;
}

// calcSurfaceDistanceCylinder, RT:3040-3088 (z axis, clipped to 0 <= z <= boxz)
__device__ bool dist_cylinder(const DSurf &s, const double *rs, const double *nd, double boxz,
                              double *lhit, OpCount &oc)
{
    const double rsc0 = s.c[0] - rs[0], rsc1 = s.c[1] - rs[1];
    double t0 = nd[0], t1 = nd[1];
    const double nrm = t0 * t0 + t1 * t1 + 0.0 * 0.0;
    const double sn = sqrt(nrm);
    t0 = t0 / sn;
    t1 = t1 / sn;
    const double rd = rsc0 * t0 + rsc1 * t1;
    const double rr = rsc0 * rsc0 + rsc1 * rsc1;
    const double d = rd * rd + s.r * s.r - rr;
    oc.flops += 20;
    if (d < 0.0) return false;
    oc.flops += 3;
    const double sq = sqrt(d);
    const double l1 = rd - sq, l2 = rd + sq;
    double l;
    if (l1 > 0.0) l = l1;
    else if (l2 > 0.0) l = l2;
    else return false;
    l = l / sqrt(nd[0] * nd[0] + nd[1] * nd[1]);
    const double zh = rs[2] + nd[2] * l;
    oc.flops += 7;
    if (zh < 0.0 || boxz < zh) return false;
    *lhit = l;
    return true;
}

// surface distance with the vertex pool read through the constant address space
__device__ inline bool surface_distance_c(const DevGeom &g, const DSurf &s, const double *p,
                                          const double *d, double *l, OpCount &oc)
{
    switch (s.kind) {
    case PT_SHAPE_POLYGON: return dist_polygon_v((cdouble *)g.vtx, s, p, d, l, oc);
    case PT_SHAPE_SPHERE: return dist_sphere(s, p, d, l, oc);
    case PT_SHAPE_CYLINDER: return dist_cylinder(s, p, d, g.box[2], l, oc);
    }
    return false;
}

__device__ inline bool surface_distance(const DevGeom &g, const DSurf &s, const double *p,
                                        const double *d, double *l, OpCount &oc)
{
    switch (s.kind) {
    case PT_SHAPE_POLYGON: return dist_polygon(g, s, p, d, l, oc);
    case PT_SHAPE_SPHERE: return dist_sphere(s, p, d, l, oc);
    case PT_SHAPE_CYLINDER: return dist_cylinder(s, p, d, g.box[2], l, oc);
    }
    return false;
}

// calcSurfaceNormal*, RT:3119-3179
__device__ void surface_normal(const DSurf &s, const double *pos, double *n, OpCount &oc)
{
    if (s.kind == PT_SHAPE_POLYGON) {
        n[0] = s.n[0]; n[1] = s.n[1]; n[2] = s.n[2];
        return;
    }
    double rcs[3] = {pos[0] - s.c[0], pos[1] - s.c[1], pos[2] - s.c[2]};
    if (s.kind == PT_SHAPE_CYLINDER) rcs[2] = 0.0;
    const double dd = dot3(rcs, rcs);
    const double inv = 1.0 / sqrt(dd);
    for (int k = 0; k < 3; ++k) n[k] = rcs[k] * inv;
    oc.flops += 14;
    if (dd > s.r * s.r)
        for (int k = 0; k < 3; ++k) n[k] = -1.0 * n[k];
}

// calcSurfaceRotation*, RT:3181-3301 (column-major 3x3)
__device__ void surface_rotation(const DevGeom &g, int idx, const DSurf &s, const double *pos,
                                 double *rot, OpCount &oc)
{
    if (s.kind == PT_SHAPE_POLYGON) {
        const double *t = g.tang + 9 * (int64_t)idx;
        for (int k = 0; k < 9; ++k) rot[k] = t[k];
        return;
    }
    double a[3] = {pos[0] - s.c[0], pos[1] - s.c[1], pos[2] - s.c[2]};
    if (s.kind == PT_SHAPE_CYLINDER) a[2] = 0.0;
    const double dd = dot3(a, a);
    const double inv = 1.0 / sqrt(dd);
    for (int k = 0; k < 3; ++k) a[k] = a[k] * inv;
    rot[6] = a[0]; rot[7] = a[1]; rot[8] = a[2];
    oc.flops += 15;
    const bool axis = (s.kind == PT_SHAPE_SPHERE) ? (fabs(a[2]) == 1.0) : (fabs(a[0]) == 1.0);
    if (axis || (a[0] * a[0] == 0.0 && a[1] * a[1] == 0.0)) {
        rot[3] = 0.0; rot[4] = 1.0; rot[5] = 0.0;
        rot[0] = 1.0; rot[1] = 0.0; rot[2] = 0.0;
    } else {
        const double r = sqrt(a[0] * a[0] + a[1] * a[1]);
        rot[3] = -a[1] / r;
        rot[4] = a[0] / r;
        rot[5] = 0.0;
        cross3(rot + 3, rot + 6, rot);
        oc.flops += 15;
    }
}

// isInsideSurface / isInsideGrain / findInsideWhichGrain, RT:1874-1985
__device__ bool inside_surface(const DSurf &s, const double *p, OpCount &oc)
{
    double d[3] = {p[0] - s.c[0], p[1] - s.c[1], p[2] - s.c[2]};
    oc.flops += 8;
    if (s.kind == PT_SHAPE_POLYGON) return dot3(d, s.n) < 0;
    if (s.kind == PT_SHAPE_SPHERE) return dot3(d, d) < s.r * s.r;
    if (s.kind == PT_SHAPE_CYLINDER) return d[0] * d[0] + d[1] * d[1] < s.r * s.r;
    return false;
}

// The grain and candidate loops below run in step on every active lane (a lane that leaves
// an inner loop early is masked), so their indices are wave-uniform and the geometry comes
// in through scalar loads.
__device__ int find_inside_which_grain(const DevGeom &g, const double *p, OpCount &oc)
{
    int count = 0, which = -1;
    for (int ig = 1; ig <= g.ngrain; ++ig) {
        bool in = true;
        // the candidate list of a grain is its internal surfaces; incident ones are
        // never internal (RT:1108-1111), so this is exactly isInsideGrain's set
        const int j0 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig]);
        const int j1 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig + 1]);
        for (int j = j0; j < j1; ++j) {
            const int idx = __builtin_amdgcn_readfirstlane(((cint *)g.cand_idx)[j]);
            if (!inside_surface(load_surf_c(g.surf, idx), p, oc)) { in = false; break; }
        }
        if (in) { which = ig; ++count; }
    }
    if (count == 1) return which;
    if (count == 0 && g.kind_grain == PT_GRAIN_BRICKS) return g.ngrain + 1;
    if (count == 0) return -1;
    return which;
}

// findNearestIncidentGrain, RT:2100-2118
__device__ int nearest_incident_grain(const DevGeom &g, const double *p, OpCount &oc)
{
    double d2min = 1.0e10;
    int best = 0;
    for (int ig = 1; ig <= g.ngrain; ++ig) {
        if (!((cint *)g.grain_incident)[ig]) continue;
        cdouble *q = (cdouble *)g.gen + 3 * ig;
        const double d0 = p[0] - q[0], d1 = p[1] - q[1], d2 = p[2] - q[2];
        const double dd = d0 * d0 + d1 * d1 + d2 * d2;
        oc.flops += 8;
        if (d2min > dd) { d2min = dd; best = ig; }
    }
    return best;
}

// findIncidentCollisionGrain, RT:2131-2163 (the last grain with a containing facet wins)
__device__ int incident_collision_grain(const DevGeom &g, const double *p, OpCount &oc)
{
    int found = -1;
    for (int ig = 1; ig <= g.ngrain; ++ig) {
        if (!((cint *)g.grain_incident)[ig]) continue;
        const int j0 = __builtin_amdgcn_readfirstlane(((cint *)g.inc_off)[ig]);
        const int j1 = __builtin_amdgcn_readfirstlane(((cint *)g.inc_off)[ig + 1]);
        for (int j = j0; j < j1; ++j) {
            const DSurf s = load_surf_c(g.surf, __builtin_amdgcn_readfirstlane(((cint *)g.inc_idx)[j]));
            if (inside_polygon((cdouble *)g.vtx + 3 * (int64_t)s.voff, s.nv, p, oc)) { found = ig; break; }
        }
    }
    return found >= 0 ? found : g.igrain_fill;
}

// Mesh grains (extension): the last grain with an incident facet containing the start
// point, -1 if none.
__device__ int mesh_start_grain(const DevGeom &g, const double *p, OpCount &oc)
{
    int found = -1;
    for (int ig = 1; ig <= g.ngrain; ++ig) {
        if (!g.grain_incident[ig]) continue;
        for (int j = g.inc_off[ig]; j < g.inc_off[ig + 1]; ++j) {
            const DSurf &s = g.surf[g.inc_idx[j]];
            if (inside_polygon(g.vtx + 3 * (int64_t)s.voff, s.nv, p, oc)) { found = ig; break; }
        }
    }
    return found;
}

// Start grain of a path of batch structure sid: the mesh rule (last grain with an
// incident facet containing the point) over the facets the structure's start locator
// lists for the point's cell; every incident facet overlapping the cell is listed.
__device__ int batch_start_grain(const DevGeom &g, const DevBatch &bt, int sid, const double *p,
                                 OpCount &oc)
{
    const int G = bt.grid;
    int iy = (int)(p[1] / g.box[1] * G), iz = (int)(p[2] / g.box[2] * G);
    iy = iy < 0 ? 0 : (iy >= G ? G - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= G ? G - 1 : iz);
    const int64_t c = ((int64_t)sid * G + iy) * G + iz;
    int found = -1;
    for (int j = bt.loc_off[c]; j < bt.loc_off[c + 1]; ++j) {
        const DSurf &s = g.surf[bt.loc_idx[j]];
        if (s.gin > found && inside_polygon(g.vtx + 3 * (int64_t)s.voff, s.nv, p, oc)) found = s.gin;
    }
    return found;
}

// State of one phonon path.
struct Path;
__device__ void scan_full(const DevGeom &g, const Path &ph, double *dmin, int *best, OpCount &oc);
struct Path {
    double pos[3];
    double dir[3];
    double fp;              // FreePath
    double mfp;             // mean free path of the path's mode (1e10: ballistic)
    double omega;           // mode frequency (Gamma/TBC laws)
    int which;              // whichV
    int cur;                // last surface hit (global index), -1 after a scattering
    int end_side;
    int flags;
    int start_grain;
    int64_t events;
    float bc[3];            // safety ball: centre and radius (br <= 0: none); no surface
    float br;               //   the scan can hit lies within it
    int sid;                // batch structure of the path
};

// ---------------------------------------------------------------------------------
// FP32 filter of the fast path.  Each candidate gets a distance interval [lo, hi] that
// contains the exact FP64 result of the reference's routine, from FP32 arithmetic on
// coordinates relative to the current position, or the verdict "unsure" (grazing, near
// an edge, tangent, near zero), which is then settled by the FP64 routine.  The event's
// surface is certified when one candidate's interval lies entirely below all others;
// its distance is then recomputed by the FP64 routine, so the trace evolves with exactly
// the FP64 numbers.  Anything else falls back to the full FP64 scan.
// Error bounds: the FP32 quantities carry a handful of roundings of relative size
// 2^-24 (conversion, products, v_rcp/v_rsq/v_sqrt at ~1 ulp); kMarg = 256 * 2^-24
// covers them with a wide safety factor.
constexpr float kEps32 = 5.9604645e-8f;
constexpr float kMarg = 256.0f * kEps32;
enum { F_MISS = 0, F_HIT = 1, F_UNSURE = 2 };

__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Axis-aligned box wall facing the ray on axis a (the box grain of setupBoxGrain,
// RT:1255-1332): cos and the sign of L are exact; the in-face test has margins.
__device__ int filt_box_axis(const DevGeom &g, int a, const double *rs, const double *nd,
                             int *widx, float *lo, float *hi)
{
    const double nda = nd[a];
    if (nda == 0.0) return F_MISS;                       // both walls back-facing
    const bool minus = nda < 0.0;
    *widx = 2 * a + (minus ? 0 : 1);
    const double P = minus ? 0.0 : g.box[a];
    const double num = P - rs[a];
    if ((num < 0.0 && !minus) || (num > 0.0 && minus)) return F_MISS;   // L < 0
    const float L = (float)num * frcp((float)nda);
    const int u = a == 2 ? 0 : a + 1, v = a == 0 ? 2 : (a == 1 ? 0 : 1);
    const float du = (float)g.box[u], dv = (float)g.box[v];
    const float ru = (float)rs[u], rv = (float)rs[v];
    const float hu = ru + (float)nd[u] * L, hv = rv + (float)nd[v] * L;
    const float mu = kMarg * (fabsf(ru) + fabsf(hu - ru) + du) + 1e-30f;
    const float mv = kMarg * (fabsf(rv) + fabsf(hv - rv) + dv) + 1e-30f;
    if (hu < -mu || hu > du + mu || hv < -mv || hv > dv + mv) return F_MISS;
    if (hu < mu || hu > du - mu || hv < mv || hv > dv - mv) return F_UNSURE;
    *lo = L * (1.0f - kMarg);
    *hi = L * (1.0f + kMarg) + 1e-30f;
    return F_HIT;
}

// Side of the start point relative to a z-cylinder in the reference's own terms: with
// rr = rsc0*rsc0 + rsc1*rsc1 (the same FP64 operations as RT:3040-3088), +1 when
// rr - r^2 > 1e-10 (rr + r^2), -1 when below minus that, else 0.  The margin exceeds
// the rounding of the reference's d = rd^2 + r^2 - rr (at most ~16 eps (rr + r^2)) by
// five orders of magnitude, so outside (+1) with rd < 0 makes its larger root
// rd + sqrt(d) negative (a miss), and inside (-1) makes its smaller root rd - sqrt(d)
// negative (the larger one is then the hit), whatever the rounding.
__device__ inline int cyl_side64(const DSurf &s, const double *rs, const double *nd, bool *leaving)
{
    const double a = s.c[0] - rs[0], b = s.c[1] - rs[1];
    const double rr = a * a + b * b, r2 = s.r * s.r;
    const double m = 1.0e-10 * (rr + r2);
    const double d = rr - r2;
    // moving away from the axis: (c - p) . d_xy < 0, with a relative margin
    const double rdu = a * nd[0] + b * nd[1];
    *leaving = rdu < -1.0e-9 * (fabs(a * nd[0]) + fabs(b * nd[1]));
    return d > m ? 1 : (d < -m ? -1 : 0);
}

// z-cylinder (calcSurfaceDistanceCylinder, RT:3040-3088)
__device__ int filt_cylinder(const DSurf &s, const double *rs, const double *nd, double boxz,
                             float *lo, float *hi)
{
    const float dx = (float)(s.c[0] - rs[0]), dy = (float)(s.c[1] - rs[1]);
    const float nx = (float)nd[0], ny = (float)nd[1];
    const float n2 = nx * nx + ny * ny;
    if (n2 < 1e-10f) return F_UNSURE;                    // near axis-parallel (0/0 in FP64)
    const float inv = frsq(n2);
    const float tx = nx * inv, ty = ny * inv;
    const float rd = dx * tx + dy * ty;
    const float rr = dx * dx + dy * dy;
    const float r = (float)s.r, r2 = r * r;
    const float rd2 = rd * rd;
    const float D = rd2 + r2 - rr;
    const float mD = kMarg * (rd2 + r2 + rr) + 1e-30f;
    if (D < -mD) return F_MISS;
    if (D <= mD) return F_UNSURE;
    const float sq = fsqrt(D);
    const float mrd = kMarg * fsqrt(rr);
    const float ml = mrd + mD * frcp(sq) * 1.0001f + kMarg * (fabsf(rd) + sq) + 1e-30f;
    const float l1 = rd - sq, l2 = rd + sq;
    float l;
    if (l1 > ml) {
        l = l1;
    } else {
        // a root at the start point (the surface just left): its sign from FP64
        bool leaving = false;
        const int side = (l1 >= -ml || (l2 >= -ml && l2 <= ml)) ? cyl_side64(s, rs, nd, &leaving) : 0;
        if (l1 >= -ml && side >= 0) return F_UNSURE;         // certified l1 < 0 only inside
        if (l2 > ml) l = l2;
        else if (l2 >= -ml) return (side > 0 && leaving) ? F_MISS : F_UNSURE;
        else return F_MISS;
    }
    const float L = l * inv;
    const float mL = ml * inv + kMarg * L;
    const float nz = (float)nd[2], rz = (float)rs[2];
    const float zh = rz + nz * L;
    const float mz = kMarg * (fabsf(rz) + fabsf(nz * L)) + fabsf(nz) * mL + 1e-30f;
    const float bz = (float)boxz;
    if (zh < -mz || zh > bz + mz) return F_MISS;
    if (zh < mz || zh > bz - mz) return F_UNSURE;
    *lo = L - mL;
    *hi = L + mL;
    return F_HIT;
}

// The reference's outside-and-leaving guard of calcSurfaceDistanceSphere (RT:2985-3038),
// with its own FP64 operations (dist_sphere's rd and rr): true when it reports no hit.
__device__ inline bool sphere_guard64(const DSurf &s, const double *rs, const double *nd)
{
    double rsc[3] = {s.c[0] - rs[0], s.c[1] - rs[1], s.c[2] - rs[2]};
    const double rd = dot3(rsc, nd);
    const double rr = dot3(rsc, rsc);
    return rr > s.r * s.r - PT_SPHERE_EPS && rd <= 0.0;
}

// sphere (calcSurfaceDistanceSphere, RT:2985-3038), including its outside-and-leaving guard
__device__ int filt_sphere(const DSurf &s, const double *rs, const double *nd, float *lo,
                           float *hi)
{
    const float dx = (float)(s.c[0] - rs[0]), dy = (float)(s.c[1] - rs[1]),
                dz = (float)(s.c[2] - rs[2]);
    const float nx = (float)nd[0], ny = (float)nd[1], nz = (float)nd[2];
    const float rd = dx * nx + dy * ny + dz * nz;
    const float rr = dx * dx + dy * dy + dz * dz;
    const float r = (float)s.r, r2 = r * r;
    const float mrr = kMarg * (rr + r2) + 1e-30f;
    const float mrd = kMarg * fsqrt(rr) + 1e-30f;
    const float out = rr - r2;                           // guard: rr > r^2 - 1e-10
    if (out > mrr) {
        if (rd < -mrd) return F_MISS;
        if (rd <= mrd) return sphere_guard64(s, rs, nd) ? F_MISS : F_UNSURE;
    } else if (out >= -mrr) {
        // on the surface (the one just left): the reference's own guard, in FP64
        return sphere_guard64(s, rs, nd) ? F_MISS : F_UNSURE;
    }
    const float rd2 = rd * rd;
    const float D = rd2 + r2 - rr;
    const float mD = kMarg * (rd2 + r2 + rr) + 1e-30f;
    if (D < -mD) return F_MISS;
    if (D <= mD) return F_UNSURE;
    const float sq = fsqrt(D);
    const float ml = mrd + mD * frcp(sq) * 1.0001f + kMarg * (fabsf(rd) + sq) + 1e-30f;
    const float l1 = rd - sq, l2 = rd + sq;
    float l;
    if (l1 > ml) l = l1;
    else if (l1 >= -ml) return F_UNSURE;
    else if (l2 > ml) l = l2;
    else if (l2 >= -ml) return F_UNSURE;
    else return F_MISS;
    *lo = l - ml;
    *hi = l + ml;
    return F_HIT;
}

// Plane of a convex polygon (calcSurfaceDistancePolygon, RT:2907-2959): F_MISS when the
// polygon certainly faces away (the reference's cosang <= 0) or its plane certainly lies
// behind the start point (L < 0); F_HIT with an interval [lo, hi] around the reference's
// FP64 plane distance L = ((v1 - p) . n) / cosang; F_UNSURE near cosang = 0 (grazing) or
// L = 0 (the facet just left).  The terms v1 - p are the reference's FP64 differences, so
// each FP32 product and sum adds a rounding of at most a few 2^-24 of the magnitudes
// involved, far inside kMarg.  Whether the hit point lies inside the polygon is settled
// for the event's winner only.
template <class NT, class VT>
__device__ inline int filt_plane(NT *n, VT *v, const double *rs, const double *nd, float *lo,
                                 float *hi)
{
    const float n0 = (float)n[0], n1 = (float)n[1], n2 = (float)n[2];
    const float c0 = (float)nd[0] * n0, c1 = (float)nd[1] * n1, c2 = (float)nd[2] * n2;
    const float c = c0 + c1 + c2;
    const float ec = kMarg * (fabsf(c0) + fabsf(c1) + fabsf(c2)) + 1e-30f;
    if (c < -ec) return F_MISS;
    if (c <= ec) return F_UNSURE;
    const float a0 = (float)(v[0] - rs[0]) * n0, a1 = (float)(v[1] - rs[1]) * n1,
                a2 = (float)(v[2] - rs[2]) * n2;
    const float a = a0 + a1 + a2;
    const float ea = kMarg * (fabsf(a0) + fabsf(a1) + fabsf(a2)) + 1e-30f;
    if (a < -ea) return F_MISS;
    if (a <= ea) return F_UNSURE;
    // relative error bound of the quotient (its own terms from v_rcp_f32, ~1 ulp, with a
    // 1e-4 relative allowance: a bound needs no correctly rounded division)
    const float rc = frcp(c);
    const float r = (ea * frcp(a) + ec * rc) * 1.0001f + kMarg;
    if (!(r < 1.0e-3f)) return F_UNSURE;
    const float L = a * rc;
    *lo = L * (1.0f - r);
    *hi = L * (1.0f + r) + 1e-30f;
    return F_HIT;
}

template <class VT>
__device__ inline int filt_polygon(VT *vtx, const DSurf &s, const double *rs, const double *nd,
                                   float *lo, float *hi)
{
    return filt_plane(s.n, vtx + 3 * (int64_t)s.voff, rs, nd, lo, hi);
}

// Running selection over candidate intervals in scan order.
// Candidate index flag: a polygon entered with its FP64 plane hit, whose in-polygon test
// is still due (surface indices stay below 2^30).
constexpr int kPlaneBit = 1 << 30;

// The intervals are FP32; an exact FP64 distance enters as the FP32 interval two units in
// the last place around it (it lies inside), and is kept for the winner.
struct Pick {
    float hi_min, lo1, lo2;
    int idx_hi, idx_lo1;
    double exact_l;       // FP64 distance of idx_hi when it was evaluated exactly
    bool exact;
    __device__ void init() { hi_min = lo1 = lo2 = 3.0e38f; idx_hi = idx_lo1 = -1; exact = false; }
    __device__ void add(int idx, float lo, float hi)
    {
        if (hi < hi_min) { hi_min = hi; idx_hi = idx; exact = false; }
        if (lo < lo1) { lo2 = lo1; lo1 = lo; idx_lo1 = idx; }
        else if (lo < lo2) lo2 = lo;
    }
    __device__ void add_exact(int idx, double l)
    {
        const float f = (float)l;
        const float e = fabsf(f) * 2.4e-7f + 1.0e-37f;
        const float lo = f - e, hi = f + e;
        if (hi < hi_min) { hi_min = hi; idx_hi = idx; exact = true; exact_l = l; }
        if (lo < lo1) { lo2 = lo1; lo1 = lo; idx_lo1 = idx; }
        else if (lo < lo2) lo2 = lo;
    }
};

// FP64 distance to box wall w of an axis-aligned box.  With n = +-e_a and v1[a] = P
// (both checked when the geometry is created) every other term of the reference's
// ((v1 - p).n) / (d.n) (RT:2928-2934) is a signed zero, so the quotient is exactly
// (P - p[a]) / d[a].  Valid where the FP32 filter certified the hit.
__device__ inline double box_wall_distance(const DevGeom &g, int w, const double *rs,
                                           const double *nd)
{
    // selects, not indexing: a dynamic index would move the path state to scratch
    const int a = w >> 1;
    const double b = a == 0 ? g.box[0] : (a == 1 ? g.box[1] : g.box[2]);
    const double r = a == 0 ? rs[0] : (a == 1 ? rs[1] : rs[2]);
    const double d = a == 0 ? nd[0] : (a == 1 ? nd[1] : nd[2]);
    const double P = (w & 1) ? b : 0.0;
    return (P - r) / d;
}

// The fast path's findCollisionSurface.  Returns false when the FP64 full scan must
// decide; else sets *dmin and *best exactly as the reference's scan would.
// The current grain's candidates are scanned group by group ("waterfall"): lanes that
// share a grain run its candidate list together with wave-uniform indices, so the
// surface records come in through scalar loads.  Curved candidates the FP32 filter
// cannot settle (mostly the surface the phonon has just left) are queued and settled
// afterwards by one FP64 evaluation site shared by all lanes, instead of one FP64 site
// per candidate of the list.
__device__ bool scan_fast(const DevGeom &g, Path &ph, double *dmin, int *best, OpCount &oc)
{
    Pick pk;
    pk.init();
    const double *rs = ph.pos, *nd = ph.dir;
    bool todo0 = true, todo1 = ph.which != 0;
    if (g.box_aligned) {
        todo0 = false;
        for (int a = 0; a < 3; ++a) {
            int w;
            float lo, hi;
            const int r = filt_box_axis(g, a, rs, nd, &w, &lo, &hi);
            oc.flops += 10;
            if (r == F_HIT) {
                pk.add(w, lo, hi);
            } else if (r == F_UNSURE) {
                double l;
                ++oc.st[256 * 6];
                if (surface_distance(g, g.surf[w], rs, nd, &l, oc)) pk.add_exact(w, l);
            }
        }
    }
    int q0 = -1, q1 = -1;                                 // queued unsure curved candidates
    if (g.lane_scan) {
        // every grain candidate is a polygon: each lane runs its own grain's list (one plane
        // record per candidate), instead of the wave taking the grains one after another
        todo0 = todo1 = false;
        const int j1 = g.cand_off[ph.which + 1];
        for (int j = g.cand_off[ph.which]; __any(j < j1); ++j) {
            if (j < j1) {
                const CandPlane c = g.cand_plane[j];
                float lo = 0.f, hi = 0.f;
                const int r = filt_plane(c.nf, c.v1, rs, nd, &lo, &hi);
                oc.flops += 12;
                double l;
                if (r == F_HIT) pk.add(c.idx | kPlaneBit, lo, hi);
                else if (r == F_UNSURE && plane_hit(c.n, c.v1, rs, nd, &l, oc))
                    pk.add_exact(c.idx | kPlaneBit, l);
            }
        }
    }
    while (__any(todo0 || todo1)) {
        if (todo0 || todo1) {
            const int mine = todo0 ? 0 : ph.which;
            const int ig = __builtin_amdgcn_readfirstlane(mine);
            if (mine == ig) {
                const int j0 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig]);
                const int j1 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig + 1]);
                for (int j = j0; j < j1; ++j) {
                    const int idx = __builtin_amdgcn_readfirstlane(((cint *)g.cand_idx)[j]);
                    const DSurf s = load_surf_c(g.surf, idx);
                    float lo = 0.f, hi = 0.f;
                    int r;
                    if (s.kind == PT_SHAPE_CYLINDER) { r = filt_cylinder(s, rs, nd, g.box[2], &lo, &hi); oc.flops += 30; }
                    else if (s.kind == PT_SHAPE_SPHERE) { r = filt_sphere(s, rs, nd, &lo, &hi); oc.flops += 30; }
                    else { r = filt_polygon((cdouble *)g.vtx, s, rs, nd, &lo, &hi); oc.flops += 12; }
                    if (r == F_HIT) {
                        // a polygon's plane interval bounds its distance from below (the hit
                        // point may still lie outside the polygon: the winner's test below)
                        pk.add(s.kind == PT_SHAPE_POLYGON ? idx | kPlaneBit : idx, lo, hi);
                    } else if (r == F_UNSURE) {
                        if (s.kind != PT_SHAPE_POLYGON && q0 < 0) {
                            q0 = idx;
                        } else if (s.kind != PT_SHAPE_POLYGON && q1 < 0) {
                            q1 = idx;
                        } else if (s.kind == PT_SHAPE_POLYGON) {
                            // the plane hit is the polygon's distance if the hit point is
                            // inside it: a lower bound that only the winner needs to turn
                            // into a hit (the in-polygon test below)
                            double l;
                            if (polygon_plane((cdouble *)g.vtx, s, rs, nd, &l, oc))
                                pk.add_exact(idx | kPlaneBit, l);
                        } else {
                            double l;
                            ++oc.st[256 * 6];
                            if (surface_distance_c(g, s, rs, nd, &l, oc)) pk.add_exact(idx, l);
                        }
                    }
                }
                if (todo0) todo0 = false;
                else todo1 = false;
            }
        }
    }
    while (__any(q0 >= 0)) {
        if (q0 >= 0) {
            double l;
            ++oc.st[256 * 6];
            if (surface_distance(g, g.surf[q0], rs, nd, &l, oc)) pk.add_exact(q0, l);
            q0 = q1;
            q1 = -1;
        }
    }
    if (pk.idx_hi < 0 || pk.idx_lo1 != pk.idx_hi || !(pk.lo2 > pk.hi_min)) return false;
    const int win = pk.idx_hi & ~kPlaneBit;
    double l = pk.exact_l;
    if (pk.idx_hi & kPlaneBit) {
        // the winning plane: its exact FP64 distance, then the in-polygon test of its hit point
        const DSurf sw = g.surf[win];
        if (!pk.exact && !polygon_plane(g.vtx, sw, rs, nd, &l, oc)) return false;
        if (!polygon_contains(g.vtx, sw, rs, nd, l, oc)) return false;
    } else if (!pk.exact) {
        if (g.box_aligned && win < 6) {
            l = box_wall_distance(g, win, rs, nd);
            oc.flops += 2;
        } else if (!surface_distance(g, g.surf[win], rs, nd, &l, oc)) {
            return false;
        }
    }
    if (!(1.0e10 > l)) return false;                      // reference: dmin starts at 1e10
    *dmin = l;
    *best = win;
    return true;
}

// findCollisionSurface, RT:1989-2096: box grain, then the current grain; the first
// strictly nearest hit wins; with no hit the previous surface is kept (*best unchanged).
// Each grain's candidate list is run by the lanes in that grain together (waterfall over
// the distinct grains of the wave), with wave-uniform, scalar-loaded records.
__device__ void scan_full(const DevGeom &g, const Path &ph, double *dmin, int *best, OpCount &oc)
{
    bool todo0 = true, todo1 = ph.which != 0;
    while (__any(todo0 || todo1)) {
        if (todo0 || todo1) {
            const int mine = todo0 ? 0 : ph.which;
            const int ig = __builtin_amdgcn_readfirstlane(mine);
            if (mine == ig) {
                const int j0 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig]);
                const int j1 = __builtin_amdgcn_readfirstlane(((cint *)g.cand_off)[ig + 1]);
                for (int j = j0; j < j1; ++j) {
                    const int idx = __builtin_amdgcn_readfirstlane(((cint *)g.cand_idx)[j]);
                    const DSurf s = load_surf_c(g.surf, idx);
                    double l;
                    if (surface_distance_c(g, s, ph.pos, ph.dir, &l, oc) && *dmin > l) {
                        *dmin = l;
                        *best = idx;
                    }
                }
                if (todo0) todo0 = false;
                else todo1 = false;
            }
        }
    }
}

// FP32 ray for BVH slab tests.  An axis the direction barely moves along (|d| <= 1e-20,
// i.e. less than 1e-10 nm over 1e10 nm) is taken as parallel: the origin must lie in
// the slab, which the 1e-3 nm box padding makes safe.
struct RayF32 {
    float o[3], inv[3], d[3];
    bool mv[3];
};

__device__ inline RayF32 ray_f32(const double *pos, const double *dir)
{
    RayF32 r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.o[k] = (float)pos[k];
        const float d = (float)dir[k];
        r.d[k] = d;
        r.mv[k] = fabsf(d) > 1.0e-20f;
        r.inv[k] = r.mv[k] ? 1.0f / d : 0.0f;
    }
    return r;
}

// Entry parameter of the ray into a padded box; false when it misses (margins 1e-5
// relative on the parameters).
__device__ inline bool slab_enter(const RayF32 &r, const float *lo, const float *hi, float *tin)
{
    float t0 = -3.0e38f, t1 = 3.0e38f;
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (r.mv[k]) {
            const float a0 = (lo[k] - r.o[k]) * r.inv[k], a1 = (hi[k] - r.o[k]) * r.inv[k];
            t0 = fmaxf(t0, fminf(a0, a1));
            t1 = fminf(t1, fmaxf(a0, a1));
        } else {
            in = in && !(r.o[k] < lo[k] || r.o[k] > hi[k]);
        }
    }
    *tin = t0;
    return in && !(t1 < -1.0e-5f * fabsf(t1) || t0 > t1 + 1.0e-5f * fabsf(t1));
}

// True unless every normal within the bounds [nlo, nhi] certainly faces away from the
// ray: the largest d . n over the bounds (each product at its larger end) below -1e-5,
// far beyond the FP32 rounding of d and of the products, so every polygon below has the
// reference's FP64 cosang < 0 (RT:2928-2930) and cannot be hit.
__device__ inline bool may_face(const RayF32 &r, const float *nlo, const float *nhi)
{
    float ub = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) ub += fmaxf(r.d[k] * nlo[k], r.d[k] * nhi[k]);
    return !(ub < -1.0e-5f);
}

// The FP64 surface tests of one BVH leaf; keeps the smallest (distance, surface index).
__device__ inline void bvh_leaf_hits(const DevGeom &g, const DevBatch &bt, const Path &ph,
                                     int first, int count, double *bl, int *bi, bool *hit,
                                     OpCount &oc)
{
    for (int k = first; k < first + count; ++k) {
        const int idx = bt.prims[k];
        const DSurf &s = g.surf[idx];
        double l;
        ++oc.tests;
        bool h;
        if (s.kind == PT_SHAPE_POLYGON) {
            // plane hit first: the in-polygon test (and its vertex loads) only for a hit
            // that would become the nearest
            h = polygon_plane(g.vtx, s, ph.pos, ph.dir, &l, oc) &&
                (l < *bl || (*hit && l == *bl && idx < *bi)) &&
                polygon_contains(g.vtx, s, ph.pos, ph.dir, l, oc);
        } else {
            h = surface_distance(g, s, ph.pos, ph.dir, &l, oc) &&
                (l < *bl || (*hit && l == *bl && idx < *bi));
        }
        if (h) {
            *bl = l;
            *bi = idx;
            *hit = true;
        }
    }
}

// Squared distance from a point to a padded box (0 inside).
__device__ inline float box_dist2(const float *lo, const float *hi, float px, float py, float pz)
{
    const float ex = fmaxf(fmaxf(lo[0] - px, px - hi[0]), 0.0f);
    const float ey = fmaxf(fmaxf(lo[1] - py, py - hi[1]), 0.0f);
    const float ez = fmaxf(fmaxf(lo[2] - pz, pz - hi[2]), 0.0f);
    return ex * ex + ey * ey + ez * ez;
}

// The LDS image of the top BVH levels sits after the lanes' stacks (stk: this lane's stack
// base, entries 256 apart).
__device__ inline TopHdr *top_hdr(int *stk)
{
    return (TopHdr *)(stk - threadIdx.x + kBvhStack * 256);
}

// Root reference of the path's grain BVH: the LDS copy when the path is bound to the
// structure staged in its workgroup's LDS (kPathTop), else the global root (-1: none).
__device__ inline int bvh_root_ref(const DevBatch &bt, const Path &ph, int *stk)
{
    if (ph.flags & kPathTop) {
        const TopHdr *h = top_hdr(stk);
        const int li = ph.which - h->g0;
        if (li >= 0 && li < h->ng) return kLdsRef + li * bt.top_k;
    }
    return ph.which >= 0 ? bt.grain_root[ph.which] : -1;
}

// A node by reference: an LDS slot or a global node (one generic load either way).
__device__ inline BvhNode bvh_node(const DevBatch &bt, int r, int *stk)
{
    const BvhNode *p = r >= kLdsRef ? &((const TopSlot *)(top_hdr(stk) + 1))[r - kLdsRef].n
                                    : bt.nodes + r;
    return *p;
}

// Wave-uniform, all lanes active: binds the lanes that have just started a path (fresh) to
// the workgroup's LDS image of the top BVH levels of their structure.  The image is
// restaged for the structure of the wave's first fresh path when no lane is bound to the
// structure it holds (the header word's count is zero): the wave claims the word (busy
// bit), copies the grains' first top_k heap slots with the child references inside the
// staged levels turned into LDS references, and publishes the new structure.  Fresh lanes
// of the staged structure then add themselves to the count; the others, and every lane
// when the count of another structure is not zero, read global nodes.  A bound lane's
// image cannot change: staging needs a zero count.
__device__ void top_bind(const DevBatch &bt, int *stk, Path &ph, bool fresh)
{
    const unsigned long long fm = __ballot(fresh);
    if (fm == 0ull) return;
    const int lane = threadIdx.x & 63;
    TopHdr *h = top_hdr(stk);
    const int s = __builtin_amdgcn_readlane(ph.sid, (int)__builtin_ctzll(fm));
    const unsigned int tag = (unsigned int)s << 16;
    int claimed = 0;
    if (lane == 0) {
        unsigned int w = __hip_atomic_load(&h->word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((w & ~0xffffu) != tag && (w & 0xffffu) == 0u && !(w & kTopBusy))
            claimed = atomicCAS(&h->word, w, kTopBusy | tag) == w;
    }
    claimed = __shfl(claimed, 0, 64);
    if (claimed) {
        const int K = bt.top_k;
        const int g0 = bt.sgrain[s];
        const int ng = min(bt.sgrain[s + 1] - g0, bt.top_cap);
        TopSlot *ls = (TopSlot *)(h + 1);
        for (int i = lane; i < ng * K * 8; i += 64) {
            const int slot = i >> 3, part = i & 7;
            const int gi = slot / K, j = slot - gi * K;
            int4 v = ((const int4 *)(bt.top + ((size_t)(g0 + gi) * kTopImage + j)))[part];
            if (part == 3) {
                // child[0], child[1]: internal children inside the staged levels -> LDS
                if (v.x >= 0 && 2 * j + 1 < K) v.x = kLdsRef + gi * K + 2 * j + 1;
                if (v.y >= 0 && 2 * j + 2 < K) v.y = kLdsRef + gi * K + 2 * j + 2;
            }
            ((int4 *)&ls[slot])[part] = v;
        }
        if (lane == 0) {
            h->g0 = g0;
            h->ng = ng;
            __hip_atomic_store(&h->word, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    const bool mine = fresh && ph.sid == s;
    const unsigned long long bm = __ballot(mine);
    int got = 0;
    if (lane == 0 && bm != 0ull) {
        unsigned int w = __hip_atomic_load(&h->word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        while (!(w & kTopBusy) && (w & ~0xffffu) == tag) {
            const unsigned int prev = atomicCAS(&h->word, w, w + (unsigned int)__popcll(bm));
            if (prev == w) {
                got = 1;
                break;
            }
            w = prev;
        }
    }
    got = __shfl(got, 0, 64);
    if (got && mine) ph.flags |= kPathTop;
}

// Leaves the LDS image (a path that ends, or a long random walk that would keep the
// image from being restaged).
__device__ inline void top_unbind(int *stk, Path &ph)
{
    atomicSub(&top_hdr(stk)->word, 1u);
    ph.flags &= ~kPathTop;
}

// findCollisionSurface (RT:1989-2096) for batch structures: the three facing box walls,
// then the current grain's candidates through its BVH.  Node tests are FP32 slab tests
// against host-padded bounds (they only prune); surfaces are tested by the reference's
// FP64 routines, and the winner is the smallest (distance, surface index): the first
// strictly nearest surface in the reference's scan order (box grain first, then the grain
// in order).  With no hit the previous surface is kept.  Nodes entered beyond `reach` are
// pruned like nodes beyond the nearest hit so far: the accelerated precisions pass the free
// path, since the event is an internal scattering whenever the nearest hit lies beyond it,
// whichever surface that is (the caller then discards the scan's surface); the FP64 kernel
// passes 1e300 and scans the whole ray, as the reference does.
__device__ void scan_bvh(const DevGeom &g, const DevBatch &bt, const Path &ph, double *dmin,
                         int *best, OpCount &oc, int *stk, double reach)
{
    double bl = 1.0e10;
    int bi = ph.cur;
    bool hit = false;
    for (int a = 0; a < 3; ++a) {
        const double d = a == 0 ? ph.dir[0] : (a == 1 ? ph.dir[1] : ph.dir[2]);
        if (d == 0.0) continue;
        const int w = 2 * a + (d < 0.0 ? 0 : 1);
        double l;
        if (surface_distance(g, g.surf[w], ph.pos, ph.dir, &l, oc) && (l < bl || (hit && l == bl && w < bi))) {
            bl = l;
            bi = w;
            hit = true;
        }
    }
    const int root = bvh_root_ref(bt, ph, stk);
    if (root >= 0) {
        const RayF32 rf = ray_f32(ph.pos, ph.dir);
        int sp = 1;
        stk[0] = root;
        while (sp > 0) {
            const BvhNode nd = bvh_node(bt, stk[--sp * 256], stk);
            ++oc.nodes;
            float t0, t1;
            bool g0 = slab_enter(rf, nd.lo[0], nd.hi[0], &t0) && may_face(rf, nd.nlo[0], nd.nhi[0]);
            bool g1 = slab_enter(rf, nd.lo[1], nd.hi[1], &t1) && may_face(rf, nd.nlo[1], nd.nhi[1]);
            // margin 1e-5 relative on the entry parameter (the padded bounds cover the rest)
            const double lim = hit && bl < reach ? bl : reach;
            g0 = g0 && !((double)t0 * (1.0 - 1.0e-5) > lim);
            g1 = g1 && !((double)t1 * (1.0 - 1.0e-5) > lim);
            // nearer child first: its hits can prune the farther one
            const bool sw = g1 && (!g0 || t1 < t0);
            const int nc = sw ? nd.child[1] : nd.child[0], fc = sw ? nd.child[0] : nd.child[1];
            const int nn = sw ? nd.count[1] : nd.count[0], fn = sw ? nd.count[0] : nd.count[1];
            const bool ng = sw ? g1 : g0;
            bool fg = sw ? g0 : g1;
            const float ft = sw ? t0 : t1;
            if (ng && nc < 0) bvh_leaf_hits(g, bt, ph, -nc - 1, nn, &bl, &bi, &hit, oc);
            fg = fg && !((double)ft * (1.0 - 1.0e-5) > (hit && bl < reach ? bl : reach));
            if (fg && fc < 0) bvh_leaf_hits(g, bt, ph, -fc - 1, fn, &bl, &bi, &hit, oc);
            if (fg && fc >= 0 && sp < kBvhStack) stk[sp++ * 256] = fc;
            if (ng && nc >= 0 && sp < kBvhStack) stk[sp++ * 256] = nc;
        }
    }
    *dmin = bl;
    *best = bi;
}

// Safety ball of a diffusive path (an acceleration with no counterpart in the
// reference).  At the current position, R = the smallest distance to any surface the
// collision scan tests (box walls and the current grain's candidates): the plane of a
// polygon (its hit points lie on that plane, RT:2928-2938), | |p - c| - r | for a
// sphere, and the radial distance for a z-cylinder.  While a free flight starts and ends
// inside the ball no candidate can be hit before it ends, so the reference's scan would
// return a distance beyond the free path and the event is an internal scattering: the
// event then skips the scan and is otherwise traced unchanged.  Distances are bounded in
// FP32 and the radius shrunk by 1e-5 relative + 1e-2 nm, far above the rounding of both
// the bound and the reference's FP64 hit distances.
// Lower bound on the distance from (px, py, pz) to one surface: the plane of a polygon,
// the shell of a sphere or cylinder (FP32; the ball's radius takes a relative and an
// absolute margin off the minimum).  V: the vertex pool's address space.
template <typename V>
__device__ inline float surf_bound(const DSurf &s, V *vtx, float px, float py, float pz)
{
    if (s.kind == PT_SHAPE_POLYGON) {
        V *v = vtx + 3 * (int64_t)s.voff;
        const float nx = (float)s.n[0], ny = (float)s.n[1], nz = (float)s.n[2];
        const float a = (px - (float)v[0]) * nx + (py - (float)v[1]) * ny + (pz - (float)v[2]) * nz;
        return fabsf(a) * __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
    }
    const float dx = px - (float)s.c[0], dy = py - (float)s.c[1];
    const float dz = s.kind == PT_SHAPE_SPHERE ? pz - (float)s.c[2] : 0.0f;
    return fabsf(__builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) - (float)s.r);
}

__device__ inline float ball_radius_of(float r) { return r * (1.0f - 1.0e-5f) - 1.0e-2f; }

// The scan order of findCollisionSurface for a path in grain `which`: the box grain's
// candidate list, then the grain's own (RT:1989-2096).  Position k of that order is the
// global surface index cand(k); n positions in all.
struct ScanOrder {
    int a0, n0, b0, n;
    __device__ ScanOrder(const DevGeom &g, int which)
    {
        a0 = g.cand_off[0];
        n0 = g.cand_off[1] - a0;
        b0 = which != 0 ? g.cand_off[which] : 0;
        n = n0 + (which != 0 ? g.cand_off[which + 1] - b0 : 0);
    }
    __device__ int cand(const DevGeom &g, int k) const
    {
        return k < n0 ? g.cand_idx[a0 + k] : g.cand_idx[b0 + k - n0];
    }
};

// The safety-ball radius of one path of a wave down to a few: the minimum of surf_bound
// over the box grain's and the path's grain's candidates, spread over the lanes (an
// exact, order-free FP32 minimum).  Uniform control flow, every lane active, arguments
// broadcast.
__device__ float lone_ball_radius(const DevGeom &g, int which, float px, float py, float pz)
{
    const int lane = threadIdx.x & 63;
    const ScanOrder so(g, which);
    float r = 3.0e38f;
    for (int c = 0; c < so.n; c += 64) {
        const int k = c + lane;
        if (k < so.n) r = fminf(r, surf_bound(g.surf[so.cand(g, k)], g.vtx, px, py, pz));
    }
    for (int m = 1; m < 64; m <<= 1) r = fminf(r, __shfl_xor(r, m, 64));
    return ball_radius_of(r);
}

// findCollisionSurface (RT:1989-2096) for one path of a wave down to a few: lane k
// evaluates the candidate at position k of the scan order with the reference's FP64
// routine, and the wave keeps the smallest distance below 1e10, the earliest position
// among equals: what scan_full's sequential strict update `dmin > l` ends with.  Uniform
// control flow, every lane active, arguments broadcast.  Returns the winner's surface
// index (-1: no hit) and its distance (1e10 with no hit).
__device__ int lone_scan(const DevGeom &g, int which, const double *pos, const double *dir,
                         double *dmin, OpCount &oc)
{
    const int lane = threadIdx.x & 63;
    const ScanOrder so(g, which);
    double bl = 1.0e10;
    int bi = -1;
    for (int c = 0; c < so.n; c += 64) {
        const int k = c + lane;
        double l = __builtin_huge_val();
        int key = 0x7fffffff, idx = -1;
        if (k < so.n) {
            idx = so.cand(g, k);
            double d;
            if (surface_distance(g, g.surf[idx], pos, dir, &d, oc) && d < bl) {
                l = d;
                key = k;
            }
        }
        for (int m = 1; m < 64; m <<= 1) {
            const double ol = __shfl_xor(l, m, 64);
            const int ok = __shfl_xor(key, m, 64);
            if (ol < l || (ol == l && ok < key)) {
                l = ol;
                key = ok;
            }
        }
        if (l < bl) {
            bl = l;
            bi = __shfl(idx, key - c, 64);
        }
    }
    *dmin = bl;
    return bi;
}

// lone_scan for geometries whose candidates are all polygons (lane_scan): lane k takes
// the plane record of position k of the scan order and the reference's FP64 plane hit
// (cosang > 0, l >= 0); the wave keeps the smallest plane distance below 1e10 at the
// earliest position.  A polygon's distance is its plane's when the hit point lies inside
// it, so if the winning plane's polygon contains its hit point, no polygon is hit nearer
// and none at the same distance earlier: that is the reference's winner.  Otherwise the
// full lone_scan decides.  One 64-byte load per lane instead of the record, vertex and
// index chain of every candidate.
__device__ int lone_scan_planes(const DevGeom &g, int which, const double *pos, const double *dir,
                                double *dmin, OpCount &oc)
{
    const int lane = threadIdx.x & 63;
    const ScanOrder so(g, which);
    double bl = 1.0e10;
    int bk = -1;
    for (int c = 0; c < so.n; c += 64) {
        const int k = c + lane;
        double l = __builtin_huge_val();
        int key = 0x7fffffff;
        if (k < so.n) {
            const CandPlane cp = g.cand_plane[k < so.n0 ? so.a0 + k : so.b0 + k - so.n0];
            double d;
            if (plane_hit(cp.n, cp.v1, pos, dir, &d, oc) && d < bl) {
                l = d;
                key = k;
            }
        }
        for (int m = 1; m < 64; m <<= 1) {
            const double ol = __shfl_xor(l, m, 64);
            const int ok = __shfl_xor(key, m, 64);
            if (ol < l || (ol == l && ok < key)) {
                l = ol;
                key = ok;
            }
        }
        if (l < bl) {
            bl = l;
            bk = key;
        }
    }
    if (bk < 0) {
        *dmin = 1.0e10;
        return -1;
    }
    const int idx = so.cand(g, bk);
    if (!polygon_contains(g.vtx, g.surf[idx], pos, dir, bl, oc)) return lone_scan(g, which, pos, dir, dmin, oc);
    *dmin = bl;
    return idx;
}

// True when both ends of the flight from p to e lie on the grain's side of every plane of
// its region by the margin: the flight then crosses no candidate's plane, and the scan
// would find no surface within it (the reference's FP64 plane distances, whose rounding
// is far below the margin, are then negative, beyond the flight, or of planes facing away).
// Wave-uniform control flow, every lane active.  Only the planes that the flights of the
// asking lanes (ask) can reach are tested: lane j's dj is plane j's signed distance from a
// point c, and both ends of a flight lie within `reach` of c, so a plane with
// dj > reach + 2m keeps both ends beyond m (the FP32 rounding of dj and reach is far below
// m).  The planes within the largest reach of the asking lanes are selected by a ballot
// and read from rp (LDS).  Returns the verdict of the asking lanes (false for the others).
__device__ inline bool in_region(const float4 *rp, float dj, bool ask, float m, float reach,
                                 const double *p, const double *e)
{
    float rmax = ask ? reach : 0.0f;
    for (int k = 1; k < 64; k <<= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, k, 64));
    unsigned long long near = __ballot(dj <= rmax + 2.0f * m);
    const float px = (float)p[0], py = (float)p[1], pz = (float)p[2];
    const float ex = (float)e[0], ey = (float)e[1], ez = (float)e[2];
    bool ok = ask;
    while (near != 0ull) {
        const int j = (int)__builtin_ctzll(near);
        near &= near - 1ull;
        const float4 q = rp[j];
        const float a = q.x * px + q.y * py + q.z * pz - q.w;
        const float b = q.x * ex + q.y * ey + q.z * ez - q.w;
        ok = ok && a > m && b > m;
    }
    return ok;
}

// Lane j's signed distance from the point c to plane pl = plane j of a region (lanes past
// the region hold a plane no point is near).
__device__ inline float region_dist(const float4 &pl, float cx, float cy, float cz)
{
    return pl.x * cx + pl.y * cy + pl.z * cz - pl.w;
}

// Lower bound on the distance from a point to the surfaces of one BVH leaf (surf_bound).
__device__ inline float bvh_leaf_bound(const DevGeom &g, const DevBatch &bt, int first, int count,
                                       float px, float py, float pz)
{
    float r = 3.0e38f;
    for (int k = first; k < first + count; ++k)
        r = fminf(r, surf_bound(g.surf[bt.prims[k]], g.vtx, px, py, pz));
    return r;
}

// The safety ball for batch structures: the box walls, then the current grain's candidates
// through its BVH, pruning nodes whose (outward-padded) box is farther than the radius
// found so far.  Leaf surfaces are bounded by surf_bound.
__device__ void ball_setup_bvh(const DevGeom &g, const DevBatch &bt, Path &ph, int *stk,
                               OpCount &oc)
{
    const float px = (float)ph.pos[0], py = (float)ph.pos[1], pz = (float)ph.pos[2];
    float r = fminf(fminf(fminf(px, (float)g.box[0] - px), fminf(py, (float)g.box[1] - py)),
                    fminf(pz, (float)g.box[2] - pz));
    r = fmaxf(r, 0.0f);
    const int root = bvh_root_ref(bt, ph, stk);
    if (root >= 0) {
        int sp = 1;
        stk[0] = root;
        while (sp > 0) {
            const BvhNode nd = bvh_node(bt, stk[--sp * 256], stk);
            ++oc.nodes;
            const float d0 = box_dist2(nd.lo[0], nd.hi[0], px, py, pz);
            const float d1 = box_dist2(nd.lo[1], nd.hi[1], px, py, pz);
            const bool sw = d1 < d0;                 // nearer child first
            const int nc = sw ? nd.child[1] : nd.child[0], fc = sw ? nd.child[0] : nd.child[1];
            const int nn = sw ? nd.count[1] : nd.count[0], fn = sw ? nd.count[0] : nd.count[1];
            const float dn = sw ? d1 : d0, df = sw ? d0 : d1;
            // a leaf's surfaces lie in its box: the distance to any of them is at least the
            // larger of the box distance and the nearest of their planes
            if (dn < r * r && nc < 0)
                r = fminf(r, fmaxf(sqrtf(dn), bvh_leaf_bound(g, bt, -nc - 1, nn, px, py, pz)));
            if (df < r * r && fc < 0)
                r = fminf(r, fmaxf(sqrtf(df), bvh_leaf_bound(g, bt, -fc - 1, fn, px, py, pz)));
            if (df < r * r && fc >= 0 && sp < kBvhStack) stk[sp++ * 256] = fc;
            if (dn < r * r && nc >= 0 && sp < kBvhStack) stk[sp++ * 256] = nc;
        }
    }
    ph.bc[0] = px;
    ph.bc[1] = py;
    ph.bc[2] = pz;
    ph.br = ball_radius_of(r);
}

// True when the free flight from the current position stays inside the safety ball.
__device__ inline bool in_ball(const Path &ph)
{
    if (!(ph.br > 0.0f)) return false;
    const float r2 = ph.br * ph.br;
    const float ax = (float)ph.pos[0] - ph.bc[0], ay = (float)ph.pos[1] - ph.bc[1],
                az = (float)ph.pos[2] - ph.bc[2];
    if (!(ax * ax + ay * ay + az * az < r2)) return false;
    if (!(ph.fp < (double)ph.br * 2.0)) return false;
    const float ex = (float)(ph.pos[0] + ph.fp * ph.dir[0]) - ph.bc[0];
    const float ey = (float)(ph.pos[1] + ph.fp * ph.dir[1]) - ph.bc[1];
    const float ez = (float)(ph.pos[2] + ph.fp * ph.dir[2]) - ph.bc[2];
    return ex * ex + ey * ey + ez * ez < r2;
}

// determinTransmit, RT:2384-2413
__device__ bool determin_transmit(const DSurf &s, int nmaterial, double omega, double max_omega,
                                  Rng &rng, double rand_in, bool present)
{
    double threshold;
    switch (s.tkind) {
    case PT_TRANS_VALUE:
        threshold = s.tparam;
        break;
    case PT_TRANS_GAMMA:
    case PT_TRANS_TBC:
        if (nmaterial == 1) {
            threshold = 1.0 / (s.tparam * omega / max_omega + 1.0);       // RT:1417-1418
            rng.oc->flops += 4;
        } else {
            const double t11 = 1.0 / (s.tparam + 1.0);                     // RT:1423
            const double om = rng.next();                                  // RT:2399
            threshold = t11 + om * (1 - t11);
            rng.oc->flops += 5;
        }
        break;
    case PT_TRANS_LORENTZ: {              // extension: 1/(1 + (omega/omega0)^2)
        const double r = omega / s.tparam;
        threshold = 1.0 / (1.0 + r * r);
        rng.oc->flops += 4;
        break;
    }
    default:
        threshold = 0.0;
    }
    const double value = present ? rand_in : rng.next();                   // RT:2410
    return value <= threshold;
}

// Kernel arithmetic variants (PT_PREC_*): the FP64 trace; the FP32 filter with FP64
// certification (bit-identical to FP64); and the fast trace, which also samples and
// reflects directions in FP32 (hardware v_sin/v_cos/v_log/v_sqrt) while positions, free
// paths, the certified hit distance and the tallies stay FP64.
constexpr int kPrecFP64 = PT_PREC_FP64;
constexpr int kPrecFast = PT_PREC_FP32_RETEST;
constexpr int kPrecFilter = PT_PREC_FP32_FILTER;

// sin and cos of 2*pi*u for u in [0, 1): V_SIN_F32/V_COS_F32 take revolutions.
__device__ inline void fsincos_rev(float u, float *sn, float *cs)
{
    *sn = __builtin_amdgcn_sinf(u);
    *cs = __builtin_amdgcn_cosf(u);
}

// Directions made in FP32 have |d| = 1 only to ~1e-7.  The sphere distance
// (RT:2985-3038) assumes |d| = 1 and its error grows as l^2 / sqrt(D): at l ~ 700 nm a
// 4e-8 error in |d|^2 puts the hit point ~5e-4 nm inside the sphere, far beyond the
// 1e-7 nm nudge, and the phonon is trapped there.  Scaling by the second-order series
// of (1 + e)^(-1/2), e = |d|^2 - 1 (exact by Sterbenz), leaves |d|^2 - 1 = O(e^3), i.e.
// |d| = 1 to FP64 rounding.
__device__ inline void renorm_dir(double *d)
{
    const double e = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - 1.0;
    const double f = 1.0 + e * (-0.5 + 0.375 * e);
    for (int k = 0; k < 3; ++k) d[k] = d[k] * f;
}

// -mfp * ln(u) with V_LOG_F32 (log2); u = 0 gives +inf as in the FP64 trace
__device__ inline double free_path32(double mfp, double u)
{
    const float l2 = __builtin_amdgcn_logf((float)u);
    return -(mfp * (double)(l2 * 0.69314718f));
}

// Surface normal and frame at the hit point in FP32 (calcSurfaceNormal*/Rotation*,
// RT:3119-3301), relative to the surface centre.
__device__ void surface_frame32(const DevGeom &g, int idx, const DSurf &s, const double *pos,
                                bool want_rot, float *n, float *rot)
{
    if (s.kind == PT_SHAPE_POLYGON) {
        n[0] = (float)s.n[0]; n[1] = (float)s.n[1]; n[2] = (float)s.n[2];
        if (want_rot) {
            const double *t = g.tang + 9 * (int64_t)idx;
            for (int k = 0; k < 9; ++k) rot[k] = (float)t[k];
        }
        return;
    }
    float a[3] = {(float)(pos[0] - s.c[0]), (float)(pos[1] - s.c[1]), (float)(pos[2] - s.c[2])};
    if (s.kind == PT_SHAPE_CYLINDER) a[2] = 0.0f;
    const float dd = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const float inv = __builtin_amdgcn_rsqf(dd);
    for (int k = 0; k < 3; ++k) a[k] = a[k] * inv;
    const float r = (float)s.r;
    const float sg = dd > r * r ? -1.0f : 1.0f;
    for (int k = 0; k < 3; ++k) n[k] = sg * a[k];
    if (!want_rot) return;
    rot[6] = a[0]; rot[7] = a[1]; rot[8] = a[2];
    const bool axis = (s.kind == PT_SHAPE_SPHERE) ? (fabsf(a[2]) == 1.0f) : (fabsf(a[0]) == 1.0f);
    if (axis || (a[0] * a[0] == 0.0f && a[1] * a[1] == 0.0f)) {
        rot[3] = 0.0f; rot[4] = 1.0f; rot[5] = 0.0f;
        rot[0] = 1.0f; rot[1] = 0.0f; rot[2] = 0.0f;
    } else {
        const float ir = __builtin_amdgcn_rsqf(a[0] * a[0] + a[1] * a[1]);
        rot[3] = -a[1] * ir;
        rot[4] = a[0] * ir;
        rot[5] = 0.0f;
        rot[0] = rot[4] * rot[8] - rot[5] * rot[7];
        rot[1] = rot[5] * rot[6] - rot[3] * rot[8];
        rot[2] = rot[3] * rot[7] - rot[4] * rot[6];
    }
}

// One iteration of the particle loop up to the event loop (RT:1632-1690).
// Returns false if the start-grain check fails (the reference EXITs the particle loop).
// lambert: the alias-mode extension draws the polar angle, cos(theta) = sqrt(u),
// after the position instead of taking the cell's fixed angle.
template <int kPrec>
__device__ bool start_path(const DevGeom &g, double cos_t, double sin_t, bool lambert, Rng &rng,
                           Path &ph, const DevBatch *bt = nullptr)
{
    OpCount &oc = *rng.oc;
    oc.flops += 7 + kFlopsSincos;
    const double y = rng.next(), z = rng.next();                            // RT:1850-1851
    ph.pos[0] = 0.0;
    ph.pos[1] = y * g.box[1] * g.scope[0];
    ph.pos[2] = z * g.box[2] * g.scope[1];
    if (kPrec == kPrecFast) {
        float ct = (float)cos_t, st = (float)sin_t;
        if (lambert) {
            const float u = (float)rng.next();
            ct = __builtin_amdgcn_sqrtf(u);
            st = __builtin_amdgcn_sqrtf(1.0f - u);
        }
        float sp, cp;
        fsincos_rev((float)rng.next(), &sp, &cp);                           // RT:1866
        ph.dir[0] = ct;
        ph.dir[1] = st * cp;
        ph.dir[2] = st * sp;
        renorm_dir(ph.dir);
    } else {
        if (lambert) {
            const double u = rng.next();
            cos_t = sqrt(u);
            sin_t = sqrt(1.0 - u);
            oc.flops += 3;
        }
        double phi = rng.next();                                            // RT:1866
        phi = phi * kTwoPi;
        double sp, cp;
        pt_sincos(phi, &sp, &cp);
        ph.dir[0] = cos_t;
        ph.dir[1] = sin_t * cp;
        ph.dir[2] = sin_t * sp;
    }
    ph.end_side = (ph.dir[0] <= 0.0) ? 1 : 0;                               // RT:1644-1652
    ph.cur = 0;                                                             // igrain=0, isurface=1
    ph.flags = 0;
    ph.events = 0;
    ph.br = -1.0f;
    int which = 0;
    if (g.start_map) {
        // a cell where both start-grain searches certainly return the same grain
        const int G = g.start_grid;
        int iy = (int)(ph.pos[1] * g.start_ih[0]), iz = (int)(ph.pos[2] * g.start_ih[1]);
        iy = iy < 0 ? 0 : (iy >= G ? G - 1 : iy);
        iz = iz < 0 ? 0 : (iz >= G ? G - 1 : iz);
        which = g.start_map[iy * G + iz];
    }
    if (which > 0) {
        ph.which = which;
        ph.start_grain = which;
    } else {
        if (g.ngrain > 0 && g.kind_grain == PT_GRAIN_VORONOI) which = nearest_incident_grain(g, ph.pos, oc);
        else if (g.ngrain > 0 && g.kind_grain == PT_GRAIN_BRICKS) which = incident_collision_grain(g, ph.pos, oc);
        else if (g.ngrain > 0 && g.kind_grain == PT_GRAIN_MESH)
            which = bt ? batch_start_grain(g, *bt, ph.sid, ph.pos, oc) : mesh_start_grain(g, ph.pos, oc);
        ph.which = which;
        ph.start_grain = which;
        if (g.kind_grain == PT_GRAIN_MESH ? which < 0
                                          : find_inside_which_grain(g, ph.pos, oc) != which) {  // RT:1673-1682
            ph.flags = PT_PATH_START_MISMATCH;
            return false;
        }
    }
    if (ph.mfp == kBallistic) ph.fp = kBallistic;                           // RT:1686-1690
    else if (kPrec == kPrecFast) ph.fp = free_path32(ph.mfp, rng.next());
    else { ph.fp = -(ph.mfp * pt_log(rng.next())); oc.flops += 1 + kFlopsLog; } // RT:2171-2172
    return true;
}

// Likelihood-ratio bookkeeping of one adjoint path.  Free flights are exponential with
// rate s_b / mfp_m, so a path with n internal scatterings over a travelled length L has
// log-likelihood n log s_b - s_b L / mfp_m + const; each specular decision on a surface of
// group g (u <= p specular, RT:2230) adds log p or log(1 - p).  The lane's group scores
// live in LDS, slot g at score[g * stride].
struct AdjLane {
    double *score;
    int stride;
    const int32_t *group;
    int nscat;
    double len;
    // per-grain tallies: the open grain's sums of dx and l (flushed, times fw = 1/mfp_m,
    // to the workgroup's accumulators gacc when the path changes grain or ends)
    double *gacc;
    int fgrain;
    double fdx, fl, fw;
    __device__ void flush()
    {
        if (gacc && fgrain >= 0) {
            atomicAdd(gacc + 2 * fgrain, fw * fdx);
            atomicAdd(gacc + 2 * fgrain + 1, fw * fl);
        }
        fgrain = -1;
        fdx = 0.0;
        fl = 0.0;
    }
    __device__ void flight(int grain, double l, double dx) { flights(grain, l, l * dx); }
    // several flights in one grain: sums of l and of l dx
    __device__ void flights(int grain, double sl, double sldx)
    {
        if (!gacc) return;
        if (grain != fgrain) {
            flush();
            fgrain = grain;
        }
        fdx += sldx;
        fl += sl;
    }
};

// createRandomDirectionAllSphere (RT:2192-2204) from its two draws: FP32 with the
// renormalisation in the fast precision, the reference's FP64 otherwise.
template <int kPrec>
__device__ inline void scatter_dir(double u_ct, double u_ph, double *dir)
{
    if (kPrec == kPrecFast) {
        const float ct = (float)u_ct * 2.0f - 1.0f;
        float sph, cph;
        fsincos_rev((float)u_ph, &sph, &cph);
        const float st = __builtin_amdgcn_sqrtf(1.0f - ct * ct);
        dir[0] = ct; dir[1] = st * cph; dir[2] = st * sph;
        renorm_dir(dir);
    } else {
        double ct = u_ct;                                                   // RT:2197
        double ph_ = u_ph;                                                  // RT:2198
        ct = ct * 2.0 - 1.0;
        ph_ = ph_ * kTwoPi;
        const double st = sqrt(1.0 - ct * ct);
        double sph, cph;
        pt_sincos(ph_, &sph, &cph);
        dir[0] = ct; dir[1] = st * cph; dir[2] = st * sph;
    }
}

// createRandomFreePath (RT:2167-2173) from its draw
template <int kPrec>
__device__ inline double free_path_of(double mfp, double u)
{
    if (kPrec == kPrecFast) return free_path32(mfp, u);
    return -(mfp * pt_log(u));                                              // RT:2171
}

// isInsideBox, RT:2177-2187
__device__ inline bool inside_box(const DevGeom &g, const double *p)
{
    return p[0] >= 0.0 - kSmallVal && p[0] <= g.box[0] + kSmallVal &&
           p[1] >= 0.0 - kSmallVal && p[1] <= g.box[1] + kSmallVal &&
           p[2] >= 0.0 - kSmallVal && p[2] <= g.box[2] + kSmallVal;
}

// One iteration of the event loop (RT:1693-1824).  Returns true when the path ended.
// kAdj: bit 1 the flights of the grain flux tallies, bit 2 the likelihood-ratio
// bookkeeping (scattering count, path length, specularity scores); 0 the plain trace.
// pre_best != -2: the wave has already run this event's scan (lone_scan: winner pre_best,
// -1 for none, at distance pre_dmin), and sets up a new ball itself (br = NaN until then).
template <int kPrec, int kAdj, bool kBatch = false>
__device__ bool step_path(const DevGeom &g, double max_omega, int64_t event_limit, Rng &rng,
                          Path &ph, AdjLane &al, const DevBatch *bt = nullptr, int *stk = nullptr,
                          int pre_best = -2, double pre_dmin = 0.0)
{
    OpCount &oc = *rng.oc;
    oc.flops += 3;                       // box + SMALL_VAL
    if (!inside_box(g, ph.pos)) {
        ph.flags |= PT_PATH_LOST;
        return true;
    }
    if (ph.events >= event_limit) {
        ph.flags |= PT_PATH_EVENT_LIMIT;
        return true;
    }
    ph.events++;

    // findCollisionSurface, RT:1989-2096: box grain, then the current grain; the first
    // strictly nearest hit wins; with no hit the previous surface is kept.
    double dmin = 1.0e10;
    int best = ph.cur;
    bool decided = false;
    // accelerated precisions: inside its safety ball the free flight ends first (dmin
    // stays beyond it and the event is an internal scattering, as the scan would find)
    const bool in_safe = kPrec != kPrecFP64 && in_ball(ph);
    if (in_safe) decided = true;
    if (!in_safe && pre_best != -2) {
        dmin = pre_dmin;
        if (pre_best >= 0) best = pre_best;
        decided = true;
    }
    if (kBatch && !in_safe && pre_best == -2) {
        scan_bvh(g, *bt, ph, &dmin, &best, oc, stk, kPrec != kPrecFP64 ? ph.fp : 1.0e300);
        decided = true;
    }
    if (!kBatch && kPrec != kPrecFP64 && !decided) {
        decided = scan_fast(g, ph, &dmin, &best, oc);
        if (!decided) {
            dmin = 1.0e10;
            best = ph.cur;
            ++oc.st[256 * 7];
        }
    }
    if (!decided) {
        scan_full(g, ph, &dmin, &best, oc);
    }
    double distance = dmin;
    ph.cur = best;

    int jreflect = 1;
    bool scattered = false;
    if (ph.fp < distance) {
        // internal scattering, RT:1735-1742; createRandomDirectionAllSphere RT:2192-2204
        distance = ph.fp;
        scattered = true;
        if (kAdj & 2) al.nscat++;
        if (kAdj & 1) al.flight(ph.which, distance, ph.dir[0]);
        for (int k = 0; k < 3; ++k) ph.pos[k] = ph.pos[k] + distance * ph.dir[k];
        oc.flops += 6;
        const double u_ct = rng.next();
        const double u_ph = rng.next();
        scatter_dir<kPrec>(u_ct, u_ph, ph.dir);
        ph.cur = -1;
        oc.flops += 8 + kFlopsSincos;    // + 6 above
    } else {
        if (kAdj & 1) al.flight(ph.which, distance, ph.dir[0]);
        for (int k = 0; k < 3; ++k) ph.pos[k] = ph.pos[k] + distance * ph.dir[k];
        oc.flops += 6;
        if (ph.cur == 0) {
            ph.end_side = 1;
        } else if (ph.cur == 1) {
            ph.end_side = 2;
        } else if (ph.cur < 0) {
            // no surface to reflect on: vgrain(0)%vsurface(0) in the reference (undefined)
            ph.flags |= PT_PATH_LOST;
            return true;
        } else if (g.periodic && ph.cur < 6 && g.surf[ph.cur].tkind == PT_TRANS_PERIODIC) {
            // extension: periodic side wall -> the opposite wall, same direction, no draw
            const int w = ph.cur;
            const bool plus = w & 1;
            if (w < 4) ph.pos[1] = plus ? ph.pos[1] - g.box[1] : ph.pos[1] + g.box[1];
            else ph.pos[2] = plus ? ph.pos[2] - g.box[2] : ph.pos[2] + g.box[2];
            ph.cur = w ^ 1;
            ph.br = -1.0f;
        } else {
            const DSurf s = g.surf[ph.cur];
            const double spec = rng.next();                                 // RT:2230
            if (kAdj & 2) {
                const int gi = al.group[ph.cur];
                if (gi >= 0) al.score[gi * al.stride] += spec <= s.p ? 1.0 / s.p : -1.0 / (1.0 - s.p);
            }
            if (spec <= s.p) {
                // reflectSpecular, RT:2275-2328
                if (determin_transmit(s, g.nmaterial, ph.omega, max_omega, rng, 0.0, false)) {
                    jreflect = 0;
                } else if (kPrec == kPrecFast) {
                    float n[3], d[3] = {(float)ph.dir[0], (float)ph.dir[1], (float)ph.dir[2]};
                    surface_frame32(g, ph.cur, s, ph.pos, false, n, nullptr);
                    float dd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
                    if (dd < 0.0f) { for (int k = 0; k < 3; ++k) n[k] = -n[k]; dd = -dd; }
                    for (int k = 0; k < 3; ++k) ph.dir[k] = d[k] - 2.0f * n[k] * dd;
                    renorm_dir(ph.dir);
                } else {
                    double n[3];
                    surface_normal(s, ph.pos, n, oc);
                    if (dot3(n, ph.dir) < 0)
                        for (int k = 0; k < 3; ++k) n[k] = -1.0 * n[k];
                    const double dd = dot3(n, ph.dir);
                    for (int k = 0; k < 3; ++k) ph.dir[k] = ph.dir[k] - 2.0 * n[k] * dd;
                    oc.flops += 19;
                }
            } else if (kPrec == kPrecFast) {
                // reflectDiffusive (RT:2332-2380) in FP32
                float n[3], rot[9], d[3] = {(float)ph.dir[0], (float)ph.dir[1], (float)ph.dir[2]};
                surface_frame32(g, ph.cur, s, ph.pos, true, n, rot);
                if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] < 0.0f)
                    for (int k = 0; k < 3; ++k) n[k] = -n[k];
                // u = sin^2(theta), RT:2213.  (float)u is 1.0f for u > 1 - 2^-25; cos(theta)
                // = 0 would leave the surface exactly tangent, and the FP32 frame's rounding
                // could then put the path on the wall's other side (inside a hole, say, where
                // it stays).  Below 1 - 2^-24, cos(theta) >= 2^-12, far above that rounding.
                const float u2 = fminf((float)rng.next(), 0.99999994f);
                const float st = __builtin_amdgcn_sqrtf(u2);
                const float ct = __builtin_amdgcn_sqrtf(1.0f - u2);
                float sph, cph;
                fsincos_rev((float)rng.next(), &sph, &cph);                 // RT:2218
                const float loc[3] = {st * cph, st * sph, ct};
                float md[3];
                for (int k = 0; k < 3; ++k)
                    md[k] = rot[k] * loc[0] + rot[3 + k] * loc[1] + rot[6 + k] * loc[2];
                const float sg1 = (d[0] * n[0] + d[1] * n[1] + d[2] * n[2] > 0.0f) ? -1.0f : 1.0f;
                for (int k = 0; k < 3; ++k) md[k] = sg1 * md[k];
                const float sg2 = (md[0] * n[0] + md[1] * n[1] + md[2] * n[2] > 0.0f) ? -1.0f : 1.0f;
                for (int k = 0; k < 3; ++k) ph.dir[k] = sg2 * md[k];
                renorm_dir(ph.dir);
                const double u = rng.next();                                // RT:2361
                if (determin_transmit(s, g.nmaterial, ph.omega, max_omega, rng, u, true)) {
                    for (int k = 0; k < 3; ++k) ph.dir[k] = -ph.dir[k];
                    jreflect = 0;
                }
            } else {
                // reflectDiffusive, RT:2332-2380
                double n[3], rot[9];
                surface_normal(s, ph.pos, n, oc);
                if (dot3(n, ph.dir) < 0.0)
                    for (int k = 0; k < 3; ++k) n[k] = -1.0 * n[k];
                surface_rotation(g, ph.cur, s, ph.pos, rot, oc);
                oc.flops += 5 + 4 + 1 + kFlopsSincos + 2 + 15 + 10;
                double st = rng.next();                                     // RT:2213
                st = sqrt(st);
                const double ct = sqrt(1.0 - st * st);
                double phd = rng.next();                                    // RT:2218
                phd = phd * kTwoPi;
                double sph, cph;
                pt_sincos(phd, &sph, &cph);
                const double loc[3] = {st * cph, st * sph, ct};
                double md[3];
                for (int k = 0; k < 3; ++k)
                    md[k] = rot[k] * loc[0] + rot[3 + k] * loc[1] + rot[6 + k] * loc[2];
                if (dot3(ph.dir, n) > 0.0)
                    for (int k = 0; k < 3; ++k) ph.dir[k] = -md[k];
                else
                    for (int k = 0; k < 3; ++k) ph.dir[k] = md[k];
                if (dot3(ph.dir, n) > 0.0)
                    for (int k = 0; k < 3; ++k) ph.dir[k] = -1.0 * ph.dir[k];
                const double u = rng.next();                                // RT:2361
                if (determin_transmit(s, g.nmaterial, ph.omega, max_omega, rng, u, true)) {
                    for (int k = 0; k < 3; ++k) ph.dir[k] = -ph.dir[k];
                    jreflect = 0;
                }
            }
        }
    }
    if (kAdj & 2) al.len += distance;
    // nudge, RT:1775-1780
    for (int k = 0; k < 3; ++k) ph.pos[k] = ph.pos[k] + PT_NUDGE * ph.dir[k];
    oc.flops += 6;
    // FreePath, RT:1783-1790
    if (ph.mfp == kBallistic) {
        ph.fp = kBallistic;
    } else {
        ph.fp = ph.fp - distance;
        oc.flops += 1;
        if (ph.fp < 1.0e-10) {
            ph.fp = free_path_of<kPrec>(ph.mfp, rng.next());
            if (kPrec != kPrecFast) oc.flops += 1 + kFlopsLog;
        }
    }
    // a diffusive path (mean free path under kBallMfpFrac of the smallest box side) that
    // scattered outside its ball asks the tail kernel (the caller that gave the scan) for a
    // new ball around the scattering point.  The persistent kernel sets up no balls: there a
    // lane inside its ball spares no work its wave does not do anyway for the other lanes'
    // scans, while setting a ball up costs a pass over the candidates (polycrystal 8%, mesh
    // batch 2.5% faster without them)
    if (kPrec != kPrecFP64) {
        const double lim = kBallMfpFrac * fmin(g.box[0], fmin(g.box[1], g.box[2]));
        if (scattered && !in_safe && ph.mfp < lim && pre_best != -2) ph.br = __builtin_nanf("");
    }
    // whichV on transmission, RT:1793-1796
    if (jreflect == 0) {
        ph.br = -1.0f;                  // the candidates change with the grain
        ph.which = g.surf[ph.cur].gout;
        if (ph.which <= 0) {            // DO ig=0,whichV,whichV with whichV=0 is undefined
            ph.flags |= PT_PATH_LOST;
            ph.end_side = 0;
            return true;
        }
    }
    return ph.end_side != 0;
}

__device__ inline double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Once the queue is empty, a wave down to this many paths hands them to the tail kernel.
constexpr int kFewPaths = 4;

// A wave's LDS buffer of the tail kernel's in-ball batches: per coordinate k and event i
// of a batch the flight fp(i) dir(i)_k and the nudge of its new direction (xn, 16 bytes,
// read by one ds_read_b128), and the positions at which the events begin (pb, 65 used).
constexpr int kChainXn = 3 * 64 * 2;
constexpr int kChainPb = 68;
// then the planes of the path's region (at most 64, one float4 each)
constexpr int kChainWords = kChainXn + 3 * kChainPb + 2 * 64;

// Makes this wave's LDS stores visible to its own later loads by other lanes.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The wave's traversal stack in the tail kernel: its lanes' BVH stacks (entries 256 apart,
// lane l of the wave at wb + l), as one array of 64 x kBvhStack entries.
constexpr int kWaveStack = 64 * kBvhStack;
__device__ inline int &wave_stack_at(int *wb, int j)
{
    return wb[(j & 63) + (j >> 6) * 256];
}

// Pushes the references of the lanes whose flag is set on the wave's stack (sp: its depth,
// wave-uniform).  False, and nothing pushed, when they would not fit.
__device__ inline bool wave_push2(int *wb, int &sp, bool p0, int r0, bool p1, int r1)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
    const int n0 = __popcll(m0), n = n0 + __popcll(m1);
    if (sp + n > kWaveStack) return false;
    const unsigned long long below = (1ull << lane) - 1ull;
    if (p0) wave_stack_at(wb, sp + __popcll(m0 & below)) = r0;
    if (p1) wave_stack_at(wb, sp + n0 + __popcll(m1 & below)) = r1;
    sp += n;
    return true;
}

// scan_bvh for a path the whole wave holds (the tail kernel): each round the lanes take up
// to 64 nodes off the wave's stack, test both children of their node (the leaves' surfaces
// by the reference's FP64 routines) and push the children to visit; the wave then keeps the
// smallest (distance, surface index) of all hits, which is what scan_bvh's nearer-child-
// first traversal ends with (its pruning and the lanes' only drop nodes that cannot hold a
// nearer hit, or an equally near one of a smaller index; nodes entered beyond `reach` as in
// scan_bvh).  The rounds follow the tree's depth, not the number of nodes visited.  False
// when the stack would overflow (the caller then runs scan_bvh).  wb: the wave's stack.
__device__ bool wave_scan_bvh(const DevGeom &g, const DevBatch &bt, const Path &ph, double *dmin,
                              int *best, OpCount &oc, int *stk, int *wb, double reach)
{
    const int lane = threadIdx.x & 63;
    double bl = 1.0e10;
    int bi = ph.cur;
    bool hit = false;
    for (int a = 0; a < 3; ++a) {
        const double d = a == 0 ? ph.dir[0] : (a == 1 ? ph.dir[1] : ph.dir[2]);
        if (d == 0.0) continue;
        const int w = 2 * a + (d < 0.0 ? 0 : 1);
        double l;
        if (surface_distance(g, g.surf[w], ph.pos, ph.dir, &l, oc) && (l < bl || (hit && l == bl && w < bi))) {
            bl = l;
            bi = w;
            hit = true;
        }
    }
    const int root = bvh_root_ref(bt, ph, stk);
    if (root >= 0) {
        const RayF32 rf = ray_f32(ph.pos, ph.dir);
        if (lane == 0) wave_stack_at(wb, 0) = root;
        int sp = 1;
        wave_lds_sync();
        while (sp > 0) {
            const int take = sp < 64 ? sp : 64;
            sp -= take;
            const int r = lane < take ? wave_stack_at(wb, sp + lane) : 0;
            wave_lds_sync();
            double lb = bl;
            int li = bi;
            bool lh = hit;
            bool p0 = false, p1 = false;
            int c0 = 0, c1 = 0;
            if (lane < take) {
                const BvhNode nd = bvh_node(bt, r, stk);
                float t0, t1;
                bool g0 = slab_enter(rf, nd.lo[0], nd.hi[0], &t0) && may_face(rf, nd.nlo[0], nd.nhi[0]);
                bool g1 = slab_enter(rf, nd.lo[1], nd.hi[1], &t1) && may_face(rf, nd.nlo[1], nd.nhi[1]);
                const double lim = hit && bl < reach ? bl : reach;
                g0 = g0 && !((double)t0 * (1.0 - 1.0e-5) > lim);
                g1 = g1 && !((double)t1 * (1.0 - 1.0e-5) > lim);
                c0 = nd.child[0];
                c1 = nd.child[1];
                if (g0 && c0 < 0) bvh_leaf_hits(g, bt, ph, -c0 - 1, nd.count[0], &lb, &li, &lh, oc);
                if (g1 && c1 < 0) bvh_leaf_hits(g, bt, ph, -c1 - 1, nd.count[1], &lb, &li, &lh, oc);
                p0 = g0 && c0 >= 0;
                p1 = g1 && c1 >= 0;
            }
            // the smallest (distance, index) over the lanes' hits: every lane ends with it
            for (int m = 1; m < 64; m <<= 1) {
                const double ol = __shfl_xor(lb, m, 64);
                const int oi = __shfl_xor(li, m, 64);
                const bool oh = __shfl_xor((int)lh, m, 64) != 0;
                if (oh && (!lh || ol < lb || (ol == lb && oi < li))) {
                    lb = ol;
                    li = oi;
                    lh = true;
                }
            }
            bl = lb;
            bi = li;
            hit = lh;
            if (!wave_push2(wb, sp, p0, c0, p1, c1)) return false;
            wave_lds_sync();
        }
    }
    *dmin = bl;
    *best = bi;
    return true;
}

// ball_setup_bvh for a path the whole wave holds, in rounds of up to 64 nodes as
// wave_scan_bvh: the radius is the minimum over the walls and the leaves of the leaf's
// bound (the larger of its box distance and its surfaces' bound), which does not depend on
// the order the leaves are met in (a leaf or subtree left out by either traversal has a box
// distance at least the radius).  False when the stack would overflow (the caller then runs
// ball_setup_bvh).
__device__ bool wave_ball_bvh(const DevGeom &g, const DevBatch &bt, Path &ph, int *stk, int *wb)
{
    const int lane = threadIdx.x & 63;
    const float px = (float)ph.pos[0], py = (float)ph.pos[1], pz = (float)ph.pos[2];
    float r = fminf(fminf(fminf(px, (float)g.box[0] - px), fminf(py, (float)g.box[1] - py)),
                    fminf(pz, (float)g.box[2] - pz));
    r = fmaxf(r, 0.0f);
    const int root = bvh_root_ref(bt, ph, stk);
    if (root >= 0) {
        if (lane == 0) wave_stack_at(wb, 0) = root;
        int sp = 1;
        wave_lds_sync();
        while (sp > 0) {
            const int take = sp < 64 ? sp : 64;
            sp -= take;
            const int ref = lane < take ? wave_stack_at(wb, sp + lane) : 0;
            wave_lds_sync();
            float lr = r;
            bool p0 = false, p1 = false;
            int c0 = 0, c1 = 0;
            if (lane < take) {
                const BvhNode nd = bvh_node(bt, ref, stk);
                const float d0 = box_dist2(nd.lo[0], nd.hi[0], px, py, pz);
                const float d1 = box_dist2(nd.lo[1], nd.hi[1], px, py, pz);
                c0 = nd.child[0];
                c1 = nd.child[1];
                if (d0 < r * r && c0 < 0)
                    lr = fminf(lr, fmaxf(sqrtf(d0), bvh_leaf_bound(g, bt, -c0 - 1, nd.count[0], px, py, pz)));
                if (d1 < r * r && c1 < 0)
                    lr = fminf(lr, fmaxf(sqrtf(d1), bvh_leaf_bound(g, bt, -c1 - 1, nd.count[1], px, py, pz)));
                p0 = d0 < r * r && c0 >= 0;
                p1 = d1 < r * r && c1 >= 0;
            }
            for (int m = 1; m < 64; m <<= 1) lr = fminf(lr, __shfl_xor(lr, m, 64));
            r = lr;
            if (!wave_push2(wb, sp, p0, c0, p1, c1)) return false;
            wave_lds_sync();
        }
    }
    ph.bc[0] = px;
    ph.bc[1] = py;
    ph.bc[2] = pz;
    ph.br = ball_radius_of(r);
    return true;
}

// A local candidate list of a path the whole wave holds (batch, tail kernel): the surfaces
// of the path's grain whose bound (surf_bound) from a centre c is at most R, gathered by one
// traversal over the leaves whose padded box lies within R of c.  Every surface of the grain
// left out lies at least R from c.  Scans and safety balls near c then test the list in one
// round instead of walking the tree again.  Returns the count, -1 when more than 64.
__device__ int wave_local_bvh(const DevGeom &g, const DevBatch &bt, const Path &ph, int *stk,
                              int *wb, int *lst, float cx, float cy, float cz, float R)
{
    const int lane = threadIdx.x & 63;
    const int root = bvh_root_ref(bt, ph, stk);
    if (root < 0) return 0;
    int n = 0;
    if (lane == 0) wave_stack_at(wb, 0) = root;
    int sp = 1;
    wave_lds_sync();
    while (sp > 0) {
        const int take = sp < 64 ? sp : 64;
        sp -= take;
        const int ref = lane < take ? wave_stack_at(wb, sp + lane) : 0;
        wave_lds_sync();
        bool p0 = false, p1 = false;
        int c0 = 0, c1 = 0, f0 = 0, n0 = 0, f1 = 0, n1 = 0;
        if (lane < take) {
            const BvhNode nd = bvh_node(bt, ref, stk);
            const bool w0 = box_dist2(nd.lo[0], nd.hi[0], cx, cy, cz) < R * R;
            const bool w1 = box_dist2(nd.lo[1], nd.hi[1], cx, cy, cz) < R * R;
            c0 = nd.child[0];
            c1 = nd.child[1];
            if (w0 && c0 < 0) { f0 = -c0 - 1; n0 = nd.count[0]; }
            if (w1 && c1 < 0) { f1 = -c1 - 1; n1 = nd.count[1]; }
            p0 = w0 && c0 >= 0;
            p1 = w1 && c1 >= 0;
        }
        // the lanes' leaf surfaces within R, one per lane and step
        int mx = n0 + n1;
        for (int m = 1; m < 64; m <<= 1) mx = max(mx, __shfl_xor(mx, m, 64));
        for (int j = 0; j < mx; ++j) {
            int idx = -1;
            if (j < n0 + n1) {
                idx = bt.prims[j < n0 ? f0 + j : f1 + (j - n0)];
                if (!(surf_bound(g.surf[idx], g.vtx, cx, cy, cz) <= R)) idx = -1;
            }
            const unsigned long long am = __ballot(idx >= 0);
            if (n + __popcll(am) > 64) return -1;
            if (idx >= 0) lst[n + __popcll(am & ((1ull << lane) - 1ull))] = idx;
            n += __popcll(am);
        }
        if (!wave_push2(wb, sp, p0, c0, p1, c1)) return -1;
        wave_lds_sync();
    }
    return n;
}

// Lower bound on the distance from a point to a polygon: the larger of the distance to its
// plane (surf_bound) and to the box of its vertices (FP32); other shapes: surf_bound.
__device__ inline float surf_bound_box(const DevGeom &g, const DSurf &s, float px, float py, float pz)
{
    const float b = surf_bound(s, g.vtx, px, py, pz);
    if (s.kind != PT_SHAPE_POLYGON) return b;
    const double *v = g.vtx + 3 * (int64_t)s.voff;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int i = 0; i < s.nv; ++i)
        for (int k = 0; k < 3; ++k) {
            const float x = (float)v[3 * i + k];
            lo[k] = fminf(lo[k], x);
            hi[k] = fmaxf(hi[k], x);
        }
    return fmaxf(b, __builtin_amdgcn_sqrtf(box_dist2(lo, hi, px, py, pz)));
}

// Lower bound on the distance from p to the box walls and the surfaces of the grain, from a
// local list of centre c and radius R (d: |p - c|, below R): the listed surfaces' bounds,
// and R - d for the others.
__device__ inline float list_ball_bound(const DevGeom &g, const int *lst, int n, float px, float py,
                                        float pz, float R, float d)
{
    const int lane = threadIdx.x & 63;
    float r = fminf(fminf(fminf(px, (float)g.box[0] - px), fminf(py, (float)g.box[1] - py)),
                    fminf(pz, (float)g.box[2] - pz));
    r = fmaxf(fminf(r, R - d), 0.0f);
    float b = lane < n ? surf_bound_box(g, g.surf[lst[lane]], px, py, pz) : 3.0e38f;
    for (int m = 1; m < 64; m <<= 1) b = fminf(b, __shfl_xor(b, m, 64));
    return fminf(r, b);
}

// scan_bvh over a local list whose region holds every point the flight can reach: the box
// walls, then the listed surfaces (one per lane, the reference's FP64 tests), keeping the
// smallest (distance, surface index).  Every surface that could be hit within the free
// path is listed, so the event is the one scan_bvh would give.
__device__ void list_scan(const DevGeom &g, const int *lst, int n, const Path &ph, double *dmin,
                          int *best, OpCount &oc)
{
    const int lane = threadIdx.x & 63;
    double bl = 1.0e10;
    int bi = ph.cur;
    bool hit = false;
    for (int a = 0; a < 3; ++a) {
        const double d = a == 0 ? ph.dir[0] : (a == 1 ? ph.dir[1] : ph.dir[2]);
        if (d == 0.0) continue;
        const int w = 2 * a + (d < 0.0 ? 0 : 1);
        double l;
        if (surface_distance(g, g.surf[w], ph.pos, ph.dir, &l, oc) && (l < bl || (hit && l == bl && w < bi))) {
            bl = l;
            bi = w;
            hit = true;
        }
    }
    double lb = bl;
    int li = bi;
    bool lh = hit;
    if (lane < n) {
        const int idx = lst[lane];
        const DSurf &s = g.surf[idx];
        double l;
        bool h;
        if (s.kind == PT_SHAPE_POLYGON) {
            h = polygon_plane(g.vtx, s, ph.pos, ph.dir, &l, oc) &&
                (l < lb || (lh && l == lb && idx < li)) &&
                polygon_contains(g.vtx, s, ph.pos, ph.dir, l, oc);
        } else {
            h = surface_distance(g, s, ph.pos, ph.dir, &l, oc) && (l < lb || (lh && l == lb && idx < li));
        }
        if (h) {
            lb = l;
            li = idx;
            lh = true;
        }
    }
    for (int m = 1; m < 64; m <<= 1) {
        const double ol = __shfl_xor(lb, m, 64);
        const int oi = __shfl_xor(li, m, 64);
        const bool oh = __shfl_xor((int)lh, m, 64) != 0;
        if (oh && (!lh || ol < lb || (ol == lb && oi < li))) {
            lb = ol;
            li = oi;
            lh = true;
        }
    }
    *dmin = lb;
    *best = li;
}

__device__ inline double pick3(const double *v, int k)
{
    return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

// In-ball events of a path that the whole wave holds (the tail kernel: every lane keeps
// the same copy of the path and computes the same values).  An in-ball event is an
// internal scattering whose new direction and next free path depend only on the path's
// three draws for it (direction: 2, free path: 1), so lane i computes those of event i of
// a batch of 64.  The positions at which the events begin come from the exact sequential
// chain p(i+1) = (p(i) + fp(i) dir(i)) + nudge dir(i+1), rounded in step_path's order:
// lane k runs coordinate k's FP64 additions over the displacements the lanes left in LDS
// (two dependent additions and one 16-byte LDS load per event), for events 0-15, then 16-63.
// Each lane then applies step_path's checks (inside the box, below the event limit,
// flight inside the ball, or kRegion: both ends of the flight inside the grain's region,
// in_region) to its event's start state, and the batch keeps the events before the first
// that fails.  Every value of the path's state is the one step_path
// would make: the path is unchanged, bit for bit, and its Rng continues at the right draw.
template <int kPrec, int kAdj, bool kRegion>
__device__ void tail_ball_run(const DevGeom &g, int64_t event_limit, Rng &rng, Path &ph,
                              AdjLane &al, double *cb, const float4 *rp = nullptr, int nr = 0,
                              float rm = 0.0f)
{
    const int lane = threadIdx.x & 63;
    double *xn = cb, *pb = cb + kChainXn;
    // regions (nr <= 64, rp: the wave's LDS copy of the planes, entries nr..63 planes no
    // point is near): lane j holds plane j for the batch balls; dl: their distances from
    // the batch's start point
    const float4 pl = kRegion ? rp[lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float dj = 3.0e38f;            // this lane's plane: distance from the batch's start point
    const uint32_t d0 = rng.draw;
    uint32_t d = d0;
    double spare = 0.0;
    int64_t nev = 0;
    while (true) {
        // draws di, di + 1, di + 2 lie in the two blocks from di / 2 on
        const uint64_t di = (uint64_t)d + 3ull * (uint64_t)lane;
        double b0[2], b1[2];
        pt_philox_block(rng.seed, rng.path, di >> 1, b0);
        pt_philox_block(rng.seed, rng.path, (di >> 1) + 1, b1);
        const bool odd = di & 1;
        const double u0 = odd ? b0[1] : b0[0];
        const double u1 = odd ? b1[0] : b0[1];
        const double u2 = odd ? b1[1] : b1[0];
        double nd[3];
        scatter_dir<kPrec>(u0, u1, nd);
        const double nfp = free_path_of<kPrec>(ph.mfp, u2);
        // lane i: the free path and direction with which event i begins (event i - 1's
        // draws; lane 0: the path's own)
        double fi = __shfl_up(nfp, 1, 64), di3[3];
        for (int k = 0; k < 3; ++k) di3[k] = __shfl_up(nd[k], 1, 64);
        if (lane == 0) {
            fi = ph.fp;
            for (int k = 0; k < 3; ++k) di3[k] = ph.dir[k];
        }
        for (int k = 0; k < 3; ++k) {
            xn[(k * 64 + lane) * 2] = fi * di3[k];
            xn[(k * 64 + lane) * 2 + 1] = PT_NUDGE * nd[k];
        }
        double p = pick3(ph.pos, lane < 3 ? lane : 2);
        if (lane < 3) pb[lane * kChainPb] = p;
        if (kRegion) {
            // the batch's safety ball: around its start point, up to the region's nearest
            // plane (the region's planes are all the scan's candidates); flights that leave
            // it take the full region test
            const float cx = (float)ph.pos[0], cy = (float)ph.pos[1], cz = (float)ph.pos[2];
            dj = region_dist(pl, cx, cy, cz);
            float r = dj;
            for (int m = 1; m < 64; m <<= 1) r = fminf(r, __shfl_xor(r, m, 64));
            ph.bc[0] = cx;
            ph.bc[1] = cy;
            ph.bc[2] = cz;
            ph.br = ball_radius_of(r);
        }
        wave_lds_sync();
        int n = 64;
        // the chain of the first 16 events, then of the other 48: a batch that ends early
        // (a flight leaving the ball and the region) pays for few additions it does not use
        for (int s0 = 0; s0 < 64; s0 = s0 == 0 ? 16 : 64) {
            const int s1 = s0 == 0 ? 16 : 64;
            if (lane < 3) {
                // eight events' displacements in flight at once, then their additions
                const double2 *x2 = (const double2 *)xn + lane * 64;
                double *po = pb + lane * kChainPb + 1;
                for (int h = s0; h < s1; h += 8) {
                    double2 v[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) v[i] = x2[h + i];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        p = (p + v[i].x) + v[i].y;
                        po[h + i] = p;
                    }
                }
            }
            wave_lds_sync();
            Path e;
            for (int k = 0; k < 3; ++k) {
                e.pos[k] = pb[k * kChainPb + lane];
                e.dir[k] = di3[k];
                e.bc[k] = ph.bc[k];
            }
            e.fp = fi;
            e.br = ph.br;
            bool safe = in_ball(e);
            if (kRegion && __any(!safe && lane >= s0 && lane < s1)) {
                double q[3];
                float reach = 0.0f;
                for (int k = 0; k < 3; ++k) q[k] = e.pos[k] + fi * di3[k];
                for (int w = 0; w < 2; ++w) {
                    const double *x = w ? q : e.pos;
                    const float ax = (float)x[0] - ph.bc[0], ay = (float)x[1] - ph.bc[1],
                                az = (float)x[2] - ph.bc[2];
                    reach = fmaxf(reach, __builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az));
                }
                // the lanes of this step only (the others hold positions of no event yet)
                const bool ask = !safe && lane >= s0 && lane < s1;
                // every lane calls it (it reduces over the wave), whatever its own verdict
                const bool inr = in_region(rp, dj, ask, rm, reach * (1.0f + 1.0e-6f), e.pos, q);
                safe = safe || inr;
            }
            const bool ok = inside_box(g, e.pos) && ph.events + lane < event_limit && safe;
            const unsigned long long okm = __ballot(ok) | (s0 == 0 ? ~0xffffull : 0xffffull);
            if (okm != ~0ull) {
                n = (int)__builtin_ctzll(~okm);
                break;
            }
        }
        if (kAdj && n > 0) {
            // the batch's n scatterings and flights for the likelihood-ratio score and the
            // grain tally, summed over the lanes (an xor butterfly: every lane ends with the
            // same sum)
            double sl = lane < n ? fi : 0.0, sx = lane < n ? fi * di3[0] : 0.0;
            for (int m = 1; m < 64; m <<= 1) {
                sl += __shfl_xor(sl, m, 64);
                sx += __shfl_xor(sx, m, 64);
            }
            if (kAdj & 1) al.flights(ph.which, sl, sx);
            if (kAdj & 2) {
                al.nscat += n;
                al.len += sl;
            }
        }
        for (int k = 0; k < 3; ++k) ph.pos[k] = pb[k * kChainPb + n];
        if (n == 64) {
            for (int k = 0; k < 3; ++k) ph.dir[k] = readlane_f64(nd[k], 63);
            ph.fp = readlane_f64(nfp, 63);                  // fp - distance = 0: resampled
        } else {
            for (int k = 0; k < 3; ++k) ph.dir[k] = readlane_f64(di3[k], n);
            ph.fp = readlane_f64(fi, n);
        }
        ph.events += n;
        nev += n;
        d += 3u * (uint32_t)n;
        if (n < 64) {
            // an odd next draw is the second half of a block: the Rng keeps it as its spare
            spare = readlane_f64(u0, n);
            break;
        }
        // the next batch overwrites the buffer: every lane has read its positions
        wave_lds_sync();
    }
    if (kRegion) ph.br = -1.0f;
    if (d != d0) {
        ph.cur = -1;
        rng.draw = d;
        if (d & 1u) rng.spare = spare;
        OpCount &oc = *rng.oc;
        // Philox blocks the path would have evaluated: one per even draw index
        oc.philox += ((d + 1u) >> 1) - ((d0 + 1u) >> 1);
        oc.flops += (uint64_t)nev * (3 + 6 + 8 + kFlopsSincos + 6 + 1 + (kPrec != kPrecFast ? 1 + kFlopsLog : 0));
    }
}

// floor(a / b) for 0 <= a < 2^53, 1 <= b < 2^53: the FP64 quotient is within one of it
__device__ inline int64_t div_floor(int64_t a, int64_t b)
{
    int64_t q = (int64_t)((double)a / (double)b);
    const int64_t r = a - q * b;
    if (r < 0) q -= 1;
    else if (r >= b) q += 1;
    return q;
}

// cell of path pid, cell = ((ik-1)*nb + (ib-1))*(nangle-1) + (ia-1) (< 2^31, checked by
// the host)
__device__ inline void cell_of_path(int64_t pid, int64_t np, int nb, int nangle, int *ib, int *ik,
                                    int *ia)
{
    uint32_t cell = (uint32_t)div_floor(pid, np);
    const uint32_t na = (uint32_t)(nangle - 1);
    *ia = (int)(cell % na) + 1;
    cell /= na;
    *ib = (int)(cell % (uint32_t)nb) + 1;
    *ik = (int)(cell / (uint32_t)nb) + 1;
}

__device__ inline double cell_omega(const DevModes &m, int ib, int ik)
{
    return m.omega ? m.omega[(ib - 1) + (int64_t)m.nb * (ik - 1)] : 0.0;
}

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Alias-mode extension: the mode of a path from two draws (Vose table).
__device__ inline int sample_mode(const DevModes &m, Rng &rng)
{
    const double u0 = rng.next(), u1 = rng.next();
    const int n = m.nb * m.nk;
    int i = (int)(u0 * (double)n);
    if (i >= n) i = n - 1;
    return u1 < m.prob[i] ? i : m.alias[i];
}

// Sets up path pid: grid mode (cell of the reference's collapsed loop) or alias mode
// (sampled mode, Lambertian incidence).  Returns start_path's verdict.
template <int kMode, int kPrec>
__device__ inline bool begin_path(const TraceParams &P, int64_t pid, Rng &rng, Path &ph, int *mode)
{
    rng.reset((uint64_t)pid);
    if (is_grid_mode(kMode)) {
        int ib, ik, ia;
        cell_of_path(pid, P.np, P.m.nb, P.nangle, &ib, &ik, &ia);
        ph.omega = cell_omega(P.m, ib, ik);
        ph.mfp = P.mfp;
        *mode = (ib - 1) + P.m.nb * (ik - 1);
        return start_path<kPrec>(P.g, P.cos_t[ia - 1], P.sin_t[ia - 1], false, rng, ph);
    } else {
        if (kMode == TRACE_BATCH) {
            // structure of the path: path_off[sid] <= pid < path_off[sid + 1]
            int lo = 0, hi = P.bt.nstruct - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (P.bt.path_off[mid] <= pid) lo = mid;
                else hi = mid - 1;
            }
            ph.sid = lo;
        }
        const int m = sample_mode(P.m, rng);
        ph.omega = P.m.omega[m];
        ph.mfp = P.m.scale ? P.m.mfp[m] / P.m.scale[m % P.m.nb] : P.m.mfp[m];
        *mode = m;
        return start_path<kPrec>(P.g, 0.0, 0.0, true, rng, ph,
                                 kMode == TRACE_BATCH ? &P.bt : nullptr);
    }
}

// The end of a path: its outcome byte (grid), per-mode tallies (alias), structure sums
// (batch), likelihood-ratio batch sums (adjoint) or grain flux flush (grains), and the
// lane's statistics.  adj_acc: the per-batch sums (the workgroup's LDS accumulators).
template <int kMode>
__device__ inline void end_path(const TraceParams &P, const Path &ph, const Rng &rng, AdjLane &al,
                                double *adj_acc, int adj_k, int mode, unsigned long long *st)
{
    const int64_t pid = (int64_t)rng.path;          // the Rng is keyed by the path's id
    constexpr bool kBatch = kMode == TRACE_BATCH;
    constexpr int kAdj = kMode == TRACE_ADJOINT || kBatch ? 2 : (kMode == TRACE_GRAINS ? 1 : 0);
    // StartSide /= EndSide (RT:1833): lost paths count, as in the reference; a path stopped
    // at the event limit has not ended and does not
    const bool transmitted = ph.end_side != 1 && !(ph.flags & PT_PATH_EVENT_LIMIT);
    if (is_grid_mode(kMode)) {
        P.outcome[pid] = transmitted ? OUT_TRANSMITTED : 0;
    } else {
        atomicAdd(P.mode_paths + mode, 1ull);
        if (transmitted) atomicAdd(P.mode_trans + mode, 1ull);
    }
    if (kBatch) {
        const double x = transmitted ? P.m.inv_mfp[mode] : 0.0;
        const int sid = ph.sid;
        double *acc = P.bt.sums + kBatchSums * sid;
        atomicAdd(acc + 0, 1.0);
        atomicAdd(acc + 1, x);
        atomicAdd(acc + 2, x * x);
        const double v = al.score[0];
        if (v != 0.0) {
            const double gs = (x - P.bt.baseline[sid]) * v;
            atomicAdd(acc + 3, gs);
            atomicAdd(acc + 4, gs * gs);
        }
        al.score[0] = 0.0;
        const unsigned n_s = (unsigned)(P.bt.path_off[sid + 1] - P.bt.path_off[sid]);
        if (atomicAdd(P.bt.done + sid, 1u) + 1u == n_s) P.bt.finish[sid] = wall_clock64();
    } else if (kAdj == 1) {
        al.flush();
    } else if (kAdj) {
        const double lam = P.m.mfp[mode];
        const int ib = mode % P.m.nb;
        const double sb = P.m.scale ? P.m.scale[ib] : 1.0;
        const double x = transmitted ? 1.0 / lam : 0.0;
        const double w = x - P.baseline;
        double *acc = adj_acc + (int)(pid % P.nbatch) * adj_k;
        atomicAdd(acc + 0, 1.0);
        atomicAdd(acc + 1, x);
        atomicAdd(acc + 2, x * x);
        for (int gi = 0; gi < P.ngroup; ++gi) {
            const double v = al.score[gi * al.stride];
            if (v != 0.0) atomicAdd(acc + 3 + gi, w * v);
            al.score[gi * al.stride] = 0.0;
        }
        atomicAdd(acc + 3 + P.ngroup + ib, w * ((double)al.nscat / sb - al.len / lam));
    }
    st[0] += 1ull;
    st[256 * 1] += (unsigned long long)ph.events;
    st[256 * 2] += rng.draw;
    if (ph.flags & PT_PATH_LOST) st[256 * 3] += 1ull;
    if (ph.flags & PT_PATH_EVENT_LIMIT) st[256 * 5] += 1ull;
}

// path flag (internal): the tail buffer was full when this path asked for a record, so it
// stays in the persistent kernel
constexpr int kPathStay = 1 << 29;

// Hands the lane's path to the tail kernel: its state, Rng position and adjoint or grain
// sums go to the next tail record, and the lane's own sums start over for its next path.
// False when the buffer is full (the path then stays).
// Tail records are written and read with device-scope accesses (sc1: through to the level
// every XCD sees), so that a tail kernel running beside the persistent one reads them
// complete: the writer waits for its stores before it stores the record's epoch.
template <class T>
__device__ inline void put_dev(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T get_dev(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// every store and load of this wave issued so far has completed
__device__ inline void wait_mem()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int kMode>
__device__ bool handoff(const TraceParams &P, const Path &ph, const Rng &rng, AdjLane &al,
                        int mode)
{
    constexpr bool kBatch = kMode == TRACE_BATCH;
    constexpr int kAdj = kMode == TRACE_ADJOINT || kBatch ? 2 : (kMode == TRACE_GRAINS ? 1 : 0);
    const bool prio = ph.mfp < P.tail_prio_mfp;
    unsigned long long slot = atomicAdd(P.queue + (prio ? kCtrTailP : kCtrTailN), 1ull);
    if (slot >= (unsigned long long)(prio ? P.tail_cap_p : P.tail_cap - P.tail_cap_p)) return false;
    if (!prio) slot += (unsigned long long)P.tail_cap_p;
    // field by field from the path's registers (a whole record assembled first would need
    // 48 more live registers at this point of the loop)
    TailRec *r = P.tail + slot;
    for (int k = 0; k < 3; ++k) {
        put_dev(&r->pos[k], ph.pos[k]);
        put_dev(&r->dir[k], ph.dir[k]);
        put_dev(&r->bc[k], ph.bc[k]);
    }
    put_dev(&r->fp, ph.fp);
    put_dev(&r->mfp, ph.mfp);
    put_dev(&r->omega, ph.omega);
    put_dev(&r->spare, rng.spare);
    put_dev(&r->len, al.len);
    put_dev(&r->fdx, al.fdx);
    put_dev(&r->fl, al.fl);
    put_dev(&r->fw, al.fw);
    put_dev(&r->pid, (int64_t)rng.path);
    put_dev(&r->events, ph.events);
    put_dev(&r->br, ph.br);
    put_dev(&r->which, ph.which);
    put_dev(&r->cur, ph.cur);
    put_dev(&r->flags, (int32_t)(ph.flags & ~kPathTop));
    put_dev(&r->mode, (int32_t)mode);
    put_dev(&r->sid, kBatch ? ph.sid : 0);
    put_dev(&r->nscat, al.nscat);
    put_dev(&r->fgrain, (kAdj & 1) ? al.fgrain : -1);
    put_dev(&r->draw, rng.draw);
    if (kAdj & 2)
        for (int gi = 0; gi < P.ngroup; ++gi) {
            put_dev(P.tail_score + slot * P.ngroup + gi, al.score[gi * al.stride]);
            al.score[gi * al.stride] = 0.0;
        }
    wait_mem();
    put_dev(&r->epoch, P.tail_epoch);
    if (kAdj & 1) {
        al.fgrain = -1;
        al.fdx = 0.0;
        al.fl = 0.0;
    }
    return true;
}

}  // namespace

// Persistent trace kernel.  Each wave takes chunks of path ids from one global queue and
// refills lanes whose path has ended (ballot + prefix count); lanes then advance one event
// per iteration.  Grid mode writes one outcome byte per path (reduced per cell later);
// alias mode adds to per-mode integer tallies (exact, order independent).  Adjoint mode
// is alias mode plus likelihood-ratio scores: each lane keeps its path's specularity
// scores in LDS, and a finished path adds (x - baseline) * score to its batch's sums in
// LDS, flushed to HBM once at the end.
// a wave-uniform 64-bit value held in scalar registers
__device__ inline int64_t uniform64(int64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Four waves per SIMD: the accelerated variants need ~140 VGPRs unconstrained (three
// waves); capped at 128 they spill a few dwords and run 14% faster on the film.
template <int kMode, bool kCount, int kPrec>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
trace_kernel(TraceParams P)
{
    constexpr bool kBatch = kMode == TRACE_BATCH;
    constexpr int kAdj = kMode == TRACE_ADJOINT || kBatch ? 2 : (kMode == TRACE_GRAINS ? 1 : 0);
    extern __shared__ double adj_lds[];
    // per-lane statistics live in LDS, not in registers held through the whole trace:
    // paths, events, draws, lost, start mismatch, event limit, FP64 retests, fallbacks
    constexpr int kStats = trace_lane_stats(kMode);
    __shared__ unsigned long long lane_stats[kStats][256];
    static_assert(sizeof(lane_stats) == trace_static_lds(kMode), "static LDS of the trace kernel");
    for (int k = 0; k < kStats; ++k) lane_stats[k][threadIdx.x] = 0ull;
    const int lane = threadIdx.x & 63;
    AdjLane al;
    al.stride = blockDim.x;
    al.score = adj_lds + threadIdx.x;
    al.group = P.surf_group;
    al.nscat = 0;
    al.len = 0.0;
    double *adj_acc = adj_lds + (kAdj ? P.ngroup * blockDim.x : 0);
    const int adj_k = kAdj ? adj_stride(P) : 0;
    al.gacc = ((kAdj & 1) && P.grain_tally) ? adj_acc + P.nbatch * adj_k : nullptr;
    al.fgrain = -1;
    al.fdx = al.fl = al.fw = 0.0;
    // batch: the lane's BVH stack (per-structure sums go straight to global atomics: a
    // path end is rare next to its events, and the LDS they would take costs a workgroup
    // per CU)
    int *stk = (int *)((char *)adj_lds + batch_lds_offset(P, blockDim.x)) + threadIdx.x;
    // batch with LDS staging: the image is empty until a wave stages a structure
    const bool top_on = kBatch && P.bt.top_k > 0;
    if (kAdj) {
        const int nlds = P.ngroup * blockDim.x + P.nbatch * adj_k + (P.grain_tally ? 2 * P.ngt : 0);
        for (int i = threadIdx.x; i < nlds; i += blockDim.x) adj_lds[i] = 0.0;
        if (kBatch && blockIdx.x == 0 && threadIdx.x == 0) *P.bt.start = wall_clock64();
        if (top_on && threadIdx.x == 0) top_hdr(stk)->word = kTopNone << 16;
        __syncthreads();
    }
    unsigned long long *st = &lane_stats[0][threadIdx.x];   // counter k at st[256 * k]
    OpCount oc = {0ull, 0ull, 0ull, 0ull, st};
    Rng rng;
    rng.oc = &oc;
    rng.seed = P.seed;
    Path ph;
    bool active = false;
    int mode = 0;
    int64_t chunk_next = 0, chunk_end = 0;       // wave-uniform slice of the queue
    bool exhausted = false;
    if (tail_mode(kMode, kPrec) && !kCount && lane == 0) atomicAdd(P.queue + kCtrMainStarted, 1ull);
    // the queue: ids below first_from, the nfirst listed paths, then the ids from first_from
    // on but those
    int64_t nfirst = -1;
    bool first_mode = false, first_done = false;   // batch: the wave's chunk is of the list
    if (P.first_ids) {
        const int64_t n = (int64_t)P.queue[kCtrFirstN];
        nfirst = n <= P.first_cap ? n : -1;
    }
    nfirst = uniform64(nfirst);
    const int64_t qtotal = P.total + (!kBatch && nfirst > 0 ? nfirst : 0);

    // long random walks go to the tail kernel (accelerated precisions; the op-counting
    // kernels keep them)
    constexpr bool kHand = tail_mode(kMode, kPrec) && !kCount;
    while (true) {
        const unsigned long long need = __ballot(!active);
        // refill in batches: start_path is long, so issue it for many lanes at once
        if (need != 0ull && !exhausted && (__popcll(need) >= P.refill_min || need == ~0ull)) {
            if (chunk_next >= chunk_end && kBatch) {
                // the listed paths first (their own queue, word 0), then the queue of this
                // wave's XCD (hardware register XCC_ID), then the others
                long long b0 = -1, e0 = -1;
                int fl = 0;
                if (lane == 0 && nfirst > 0 && !first_done) {
                    const unsigned long long b = atomicAdd(P.queue, (unsigned long long)P.chunk);
                    if ((int64_t)b < nfirst) {
                        b0 = (long long)b;
                        e0 = b0 + P.chunk < nfirst ? b0 + P.chunk : nfirst;
                        fl = 1;
                    }
                }
                if (lane == 0 && !fl) {
                    const int home = __builtin_amdgcn_s_getreg(20 | (3 << 11)) & (kXcdQueues - 1);
                    for (int t = 0; t < kXcdQueues; ++t) {
                        const int q = (home + t) & (kXcdQueues - 1);
                        const long long qs = P.bt.qoff[q], qe = P.bt.qoff[q + 1];
                        if (qs >= qe) continue;
                        const unsigned long long b = atomicAdd(P.bt.qctr + q, (unsigned long long)P.chunk);
                        if ((long long)b < qe - qs) {
                            b0 = qs + (long long)b;
                            e0 = b0 + P.chunk < qe ? b0 + P.chunk : qe;
                            break;
                        }
                    }
                }
                b0 = __shfl(b0, 0, 64);
                e0 = __shfl(e0, 0, 64);
                first_mode = __builtin_amdgcn_readfirstlane(fl) != 0;
                if (!first_mode) first_done = true;
                chunk_next = uniform64(b0 < 0 ? P.total : b0);
                chunk_end = uniform64(e0 < 0 ? P.total : e0);
                if (b0 < 0) exhausted = true;
            } else if (chunk_next >= chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(P.queue, (unsigned long long)P.chunk);
                base = __shfl(base, 0, 64);
                chunk_next = uniform64((int64_t)base);
                chunk_end = chunk_next + P.chunk < qtotal ? chunk_next + P.chunk : qtotal;
                if (chunk_next >= qtotal) exhausted = true;
            }
            if (!exhausted) {
                const int64_t avail = chunk_end - chunk_next;
                const int rank = __popcll(need & ((1ull << lane) - 1ull));
                const int nneed = __popcll(need);
                bool fresh = false;
                if (!active && rank < avail) {
                    const int64_t v = chunk_next + rank;
                    const int64_t a = P.first_from;
                    const int64_t pid = kBatch ? (first_mode ? P.first_ids[v] : v)
                                               : (nfirst <= 0 || v < a ? v
                                                                       : (v < a + nfirst ? P.first_ids[v - a] : v - nfirst));
                    al.nscat = 0;
                    al.len = 0.0;
                    const bool started = begin_path<kMode, kPrec>(P, pid, rng, ph, &mode);
                    if (nfirst >= 0 && (kBatch ? !first_mode : v >= a + nfirst) && ph.mfp < P.first_mfp) {
                        // a listed path: it ran from the list
                    } else if (started) {
                        active = true;
                        fresh = true;
                        if (kAdj & 1) al.fw = P.m.inv_mfp[mode];
                    } else {
                        if (is_grid_mode(kMode)) P.outcome[pid] = OUT_MISMATCH;
                        st[0] += 1ull;
                        st[256 * 4] += 1ull;
                        st[256 * 2] += rng.draw;
                    }
                }
                chunk_next += (avail < nneed ? avail : nneed);
                if (top_on) top_bind(P.bt, stk, ph, fresh);
            }
        }
        if (!__any(active)) {
            if (exhausted) break;
            continue;
        }
        if (top_on && active && (ph.flags & kPathTop) && ph.events >= kTopUnbindEvents)
            top_unbind(stk, ph);
        if (kHand) {
            // a path past tail_events events (while the queue lasts, kTailLongFactor times as
            // many above the priority mean free path), or one of the last few of the wave once
            // the queue is empty, goes to the tail kernel
            const bool few = exhausted && __popcll(__ballot(active)) <= kFewPaths;
            // (batch: an iteration of the persistent kernel is ~50 times longer than on the
            // Voronoi cells, so the paths of the queue's list that are already 20 mean free
            // paths deep, the ones likely to walk long, go after an eighth of tail_events)
            const int64_t hand_at = ph.mfp < P.tail_prio_mfp
                                        ? (kBatch && ph.mfp < P.first_mfp && ph.pos[0] > 20.0 * ph.mfp
                                               ? (P.tail_events + 7) / 8 : P.tail_events)
                                        : (exhausted ? P.tail_events : kTailLongFactor * P.tail_events);
            if (active && ph.end_side == 0 && !(ph.flags & kPathStay) && (few || ph.events >= hand_at)) {
                if (handoff<kMode>(P, ph, rng, al, mode)) {
                    if (top_on && (ph.flags & kPathTop)) top_unbind(stk, ph);
                    active = false;
                } else {
                    ph.flags |= kPathStay;
                }
            }
        }
        if (active) {
            if (ph.end_side != 0 ||
                step_path<kPrec, kAdj, kBatch>(P.g, P.m.max_omega, P.event_limit, rng, ph, al,
                                               &P.bt, stk)) {
                end_path<kMode>(P, ph, rng, al, adj_acc, adj_k, mode, st);
                if (top_on && (ph.flags & kPathTop)) top_unbind(stk, ph);
                active = false;
            }
        }
    }
    // this wave hands nothing over any more
    if (kHand && lane == 0) atomicAdd(P.queue + kCtrMainDone, 1ull);
    if (kAdj) {
        __syncthreads();
        for (int i = threadIdx.x; i < P.nbatch * adj_k; i += blockDim.x)
            if (adj_acc[i] != 0.0) atomicAdd(P.adj + i, adj_acc[i]);
        if (al.gacc)
            for (int i = threadIdx.x; i < 2 * P.ngt; i += blockDim.x)
                if (al.gacc[i] != 0.0) atomicAdd(P.grain_tally + i, al.gacc[i]);
    }
    // stats slots: paths, events, draws, lost, mismatch, limit -> stats[0..5];
    // retests, fallbacks -> stats[10], stats[11]
    for (int k = 0; k < kStats; ++k) {
        const unsigned long long v = wave_sum(st[256 * k]);
        if (lane == 0) atomicAdd(P.stats + (k < 6 ? k : k + 4), v);
    }
    if (kCount) {
        const unsigned long long f = wave_sum(oc.flops), b = wave_sum(oc.philox);
        const unsigned long long bn = wave_sum(oc.nodes), bs = wave_sum(oc.tests);
        if (lane == 0) {
            atomicAdd(P.stats + 7, f);
            atomicAdd(P.stats + 8, b);
            atomicAdd(P.stats + 12, bn);
            atomicAdd(P.stats + 13, bs);
        }
    }
}

// The pre-pass of the queue's list: the ids from first_from on of the paths whose mean free
// path is below first_mfp (the mode each samples, as begin_path does), in no particular
// order; at most first_cap of them, counted in full.
__global__ void __launch_bounds__(256) first_scan_kernel(TraceParams P)
{
    // four paths per lane at once (their draws, then their table loads), so the wave waits
    // for one round of loads per four paths
    constexpr int kUnroll = 4;
    const int lane = threadIdx.x & 63;
    if (P.live_gate > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        // the persistent kernel, which follows, must not take the room of the live tail
        // workgroups before they are resident: wait for them, at most live_ticks / 4000
        const uint64_t t_end = wall_clock64() + (uint64_t)P.live_ticks / 4000u;
        while (get_dev(P.queue + kCtrLiveStarted) < (unsigned long long)P.live_gate && wall_clock64() < t_end)
            __builtin_amdgcn_s_sleep(8);
    }
    Rng rng;
    OpCount oc = {0ull, 0ull, 0ull, 0ull, nullptr};
    rng.oc = &oc;
    rng.seed = P.seed;
    const int n = P.m.nb * P.m.nk;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b = P.first_from + (int64_t)blockIdx.x * blockDim.x; b < P.total; b += kUnroll * stride) {
        int col[kUnroll];
        double u1[kUnroll];
        for (int k = 0; k < kUnroll; ++k) {
            const int64_t pid = b + k * stride + threadIdx.x;
            // sample_mode's two draws and column
            rng.reset((uint64_t)(pid < P.total ? pid : 0));
            const double u0 = rng.next();
            u1[k] = rng.next();
            const int i = (int)(u0 * (double)n);
            col[k] = i < n ? i : n - 1;
        }
        for (int k = 0; k < kUnroll; ++k) {
            const int64_t pid = b + k * stride + threadIdx.x;
            const int m = u1[k] < P.m.prob[col[k]] ? col[k] : P.m.alias[col[k]];
            const double mfp = P.m.scale ? P.m.mfp[m] / P.m.scale[m % P.m.nb] : P.m.mfp[m];
            const bool hit = pid < P.total && mfp < P.first_mfp;
            const unsigned long long hm = __ballot(hit);
            if (hm == 0ull) continue;
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(P.queue + kCtrFirstN, (unsigned long long)__popcll(hm));
            base = __shfl(base, 0, 64);
            const int64_t slot = (int64_t)base + __popcll(hm & ((1ull << lane) - 1ull));
            if (hit && slot < P.first_cap) P.first_ids[slot] = pid;
        }
    }
}

// The live tail kernel's next record (lane 0): the oldest complete record of the priority
// part, else of the general part, taken by a compare-and-swap on that part's next index, so
// that it and the tail kernel after the persistent one (atomicAdd on the same index) never
// take the same record.  Polls, sleeping, until the persistent kernel's waves have all
// ended or the wall clock passes t_end (t_wait if not all of them have begun); -1: nothing
// more for this wave (the tail kernel after the persistent one takes what is left).
__device__ int64_t live_claim(const TraceParams &P, uint64_t t_wait, uint64_t t_end)
{
    unsigned long long *q = P.queue;
    while (true) {
        const bool done = get_dev(q + kCtrMainDone) >= (unsigned long long)P.main_waves;
        for (int part = 0; part < 2; ++part) {
            const int64_t cap = part ? P.tail_cap - P.tail_cap_p : P.tail_cap_p;
            const int64_t n = min((int64_t)get_dev(q + (part ? kCtrTailN : kCtrTailP)), cap);
            unsigned long long *next = q + (part ? kCtrTailNextN : kCtrTailNext);
            int64_t t = (int64_t)get_dev(next);
            while (t < n) {
                const int64_t slot = part ? P.tail_cap_p + t : t;
                if (get_dev(&P.tail[slot].epoch) != P.tail_epoch) break;   // being written
                const unsigned long long seen =
                    atomicCAS(next, (unsigned long long)t, (unsigned long long)t + 1ull);
                if ((int64_t)seen == t) return slot;
                t = (int64_t)seen;
            }
        }
        const uint64_t now = wall_clock64();
        if (done || now > t_end ||
            (now > t_wait && get_dev(q + kCtrMainStarted) < (unsigned long long)P.main_waves))
            return -1;
        __builtin_amdgcn_s_sleep(32);
    }
}

// Tail kernel: the long random walks the persistent kernel handed over (paths past
// tail_events events, and the last few of each wave once its queue ran dry), one path per
// wave.  Every lane of the wave keeps the same copy of the path and computes the same
// values, so the control flow is uniform and the whole wave is at the path's service:
// its ball set-up and its scans spread the candidates over the lanes (lone_ball_radius,
// lone_scan; mesh grains: the wave's local candidate list, else BVH traversals of up to 64
// nodes a round), its in-ball events run 64 at a time (tail_ball_run), and the rest of
// each event runs in step_path.  Launched beside the persistent kernel (tail_live), it
// takes records as they are handed over (live_claim); launched after it, it takes what is
// left.  Lane 0 ends the path (end_path: the same tallies as the
// persistent kernel; adjoint batch sums and grain sums go to HBM directly) and holds the
// statistics.  The path takes the same steps as it would have in the persistent kernel.
template <int kMode, int kPrec>
__global__ void __launch_bounds__(256) tail_kernel(TraceParams P)
{
    constexpr bool kBatch = kMode == TRACE_BATCH;
    constexpr int kAdj = kMode == TRACE_ADJOINT || kBatch ? 2 : (kMode == TRACE_GRAINS ? 1 : 0);
    constexpr int kStats = trace_lane_stats(kMode);
    __shared__ unsigned long long lane_stats[kStats][256];
    __shared__ __attribute__((aligned(16))) double chain[4][kChainWords];
    __shared__ double tscore[kAdjMaxGroups * 4];          // [group * 4 + wave]
    __shared__ int tstk[kBatch ? kBvhStack * 256 : 1];   // BVH stacks, entries 256 apart
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < kStats; ++k) lane_stats[k][threadIdx.x] = 0ull;
    unsigned long long *st = &lane_stats[0][threadIdx.x];
    OpCount oc = {0ull, 0ull, 0ull, 0ull, st};
    Rng rng;
    rng.oc = &oc;
    rng.seed = P.seed;
    AdjLane al;
    al.stride = 4;
    al.score = tscore + wave;
    al.group = P.surf_group;
    al.gacc = ((kAdj & 1) && P.grain_tally && lane == 0) ? P.grain_tally : nullptr;
    const int adj_k = kAdj ? adj_stride(P) : 0;
    int *stk = tstk + (kBatch ? threadIdx.x : 0);
    int *wb = tstk + (kBatch ? (threadIdx.x & ~63) : 0);    // the wave's stack (batch)
    double *cb = chain[wave];
    float4 *rl = (float4 *)(cb + kChainXn + 3 * kChainPb);   // the staged region's planes
    int *lst = (int *)rl;                 // batch (no regions): the wave's local list, 64 entries
    int reg_grain = -1;
    const double lim = kBallMfpFrac * fmin(P.g.box[0], fmin(P.g.box[1], P.g.box[2]));
    // the priority part of the buffer first, then the general part; the live kernel (beside
    // the persistent one) takes complete records as they come
    const int64_t np = (int64_t)min(P.queue[kCtrTailP], (unsigned long long)P.tail_cap_p);
    const int64_t ng = (int64_t)min(P.queue[kCtrTailN], (unsigned long long)(P.tail_cap - P.tail_cap_p));
    const uint64_t t0 = P.tail_live ? wall_clock64() : 0ull;
    if (P.tail_live && threadIdx.x == 0) atomicAdd(P.queue + kCtrLiveStarted, 1ull);
    const uint64_t t_wait = t0 + (uint64_t)P.live_ticks / 200u, t_end = t0 + (uint64_t)P.live_ticks;
    while (true) {
        int64_t slot = -1;
        if (lane == 0) {
            if (P.tail_live) {
                __builtin_amdgcn_s_setprio(0);
                slot = live_claim(P, t_wait, t_end);
                __builtin_amdgcn_s_setprio(3);
            } else {
                const int64_t t = (int64_t)atomicAdd(P.queue + kCtrTailNext, 1ull);
                if (t < np) {
                    slot = t;
                } else {
                    const int64_t u = (int64_t)atomicAdd(P.queue + kCtrTailNextN, 1ull);
                    if (u < ng) slot = P.tail_cap_p + u;
                }
            }
        }
        slot = uniform64(__shfl(slot, 0, 64));
        if (slot < 0) break;
        const TailRec *r = P.tail + slot;
        Path ph;
        for (int k = 0; k < 3; ++k) {
            ph.pos[k] = get_dev(&r->pos[k]);
            ph.dir[k] = get_dev(&r->dir[k]);
            ph.bc[k] = get_dev(&r->bc[k]);
        }
        ph.fp = get_dev(&r->fp);
        ph.mfp = get_dev(&r->mfp);
        ph.omega = get_dev(&r->omega);
        ph.which = get_dev(&r->which);
        ph.cur = get_dev(&r->cur);
        ph.end_side = 0;
        ph.flags = get_dev(&r->flags);
        ph.start_grain = -1;
        ph.events = get_dev(&r->events);
        ph.br = get_dev(&r->br);
        ph.sid = get_dev(&r->sid);
        const int rmode = get_dev(&r->mode);
        rng.path = (uint64_t)get_dev(&r->pid);
        rng.draw = get_dev(&r->draw);
        rng.spare = get_dev(&r->spare);
        al.nscat = get_dev(&r->nscat);
        al.len = get_dev(&r->len);
        al.fgrain = get_dev(&r->fgrain);
        al.fdx = get_dev(&r->fdx);
        al.fl = get_dev(&r->fl);
        al.fw = get_dev(&r->fw);
        if (kAdj & 2)
            for (int gi = 0; gi < P.ngroup; ++gi)
                al.score[gi * al.stride] = get_dev(P.tail_score + slot * P.ngroup + gi);
        // batch: the wave's local candidate list (none yet)
        int lc_grain = -1, lc_n = 0;
        float lc_x = 0.0f, lc_y = 0.0f, lc_z = 0.0f, lc_r = 0.0f;
        while (true) {
            int pre_best = -2;
            double pre_dmin = 0.0;
            if (!kBatch && P.g.lane_scan && ph.mfp < lim) {
                // every candidate is a polygon: the grain's region replaces the safety balls
                ph.br = -1.0f;
                const int r0 = P.g.region_off[ph.which], nr = P.g.region_off[ph.which + 1] - r0;
                if (nr > 0 && nr <= 64) {
                    if (reg_grain != ph.which) {
                        // the wave's LDS copy of the region, padded with planes no point is near
                        rl[lane] = lane < nr ? P.g.region[r0 + lane] : make_float4(0.0f, 0.0f, 0.0f, -3.0e38f);
                        wave_lds_sync();
                        reg_grain = ph.which;
                    }
                    double q[3];
                    for (int k = 0; k < 3; ++k) q[k] = ph.pos[k] + ph.fp * ph.dir[k];
                    const float dj = region_dist(rl[lane], (float)ph.pos[0], (float)ph.pos[1],
                                                 (float)ph.pos[2]);
                    if (in_region(rl, dj, true, P.g.region_margin, (float)ph.fp * (1.0f + 1.0e-6f),
                                  ph.pos, q))
                        tail_ball_run<kPrec, kAdj, true>(P.g, P.event_limit, rng, ph, al, cb, rl, nr,
                                                         P.g.region_margin);
                }
            } else {
                if (__builtin_isnan(ph.br)) {
                    // the ball step_path asked for at its last scattering
                    if (kBatch) {
                        const float px = (float)ph.pos[0], py = (float)ph.pos[1], pz = (float)ph.pos[2];
                        float d = 3.0e38f;
                        if (lc_grain == ph.which) {
                            const float dx = px - lc_x, dy = py - lc_y, dz = pz - lc_z;
                            d = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                        }
                        if (!(d <= 0.5f * lc_r)) {
                            // a new list around this point (a smaller one if it would hold more
                            // than 64 surfaces)
                            lc_grain = -1;
                            float R = fmaxf(kLocalMfps * (float)ph.mfp, kLocalMin);
                            for (int t = 0; t < 3 && lc_grain < 0; ++t, R *= 0.25f) {
                                const int n = wave_local_bvh(P.g, P.bt, ph, stk, wb, lst, px, py, pz, R);
                                if (n >= 0) {
                                    lc_grain = ph.which;
                                    lc_n = n;
                                    lc_x = px;
                                    lc_y = py;
                                    lc_z = pz;
                                    lc_r = R;
                                    d = 0.0f;
                                }
                            }
                            wave_lds_sync();
                        }
                        if (lc_grain == ph.which) {
                            ph.bc[0] = px;
                            ph.bc[1] = py;
                            ph.bc[2] = pz;
                            ph.br = ball_radius_of(list_ball_bound(P.g, lst, lc_n, px, py, pz, lc_r, d));
                        } else if (!wave_ball_bvh(P.g, P.bt, ph, stk, wb)) {
                            ball_setup_bvh(P.g, P.bt, ph, stk, oc);
                        }
                    } else {
                        for (int k = 0; k < 3; ++k) ph.bc[k] = (float)ph.pos[k];
                        ph.br = lone_ball_radius(P.g, ph.which, ph.bc[0], ph.bc[1], ph.bc[2]);
                    }
                }
                if (in_ball(ph)) tail_ball_run<kPrec, kAdj, false>(P.g, P.event_limit, rng, ph, al, cb);
            }
            if (kBatch && !in_ball(ph)) {
                double d;
                int b;
                bool near = false;
                if (lc_grain == ph.which) {
                    // the flight stays inside the list's region (the margins of a safety ball)
                    const float dx = (float)ph.pos[0] - lc_x, dy = (float)ph.pos[1] - lc_y,
                                dz = (float)ph.pos[2] - lc_z;
                    near = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) + (float)ph.fp * (1.0f + 1.0e-6f) <
                           ball_radius_of(lc_r);
                }
                if (near) {
                    list_scan(P.g, lst, lc_n, ph, &d, &b, oc);
                    pre_best = b;
                    pre_dmin = d;
                } else if (wave_scan_bvh(P.g, P.bt, ph, &d, &b, oc, stk, wb, ph.fp)) {
                    pre_best = b;
                    pre_dmin = d;
                }
            }
            if (!kBatch && !in_ball(ph))
                pre_best = P.g.lane_scan ? lone_scan_planes(P.g, ph.which, ph.pos, ph.dir, &pre_dmin, oc)
                                         : lone_scan(P.g, ph.which, ph.pos, ph.dir, &pre_dmin, oc);
            if (step_path<kPrec, kAdj, kBatch>(P.g, P.m.max_omega, P.event_limit, rng, ph, al,
                                               &P.bt, stk, pre_best, pre_dmin))
                break;
        }
        if (lane == 0) end_path<kMode>(P, ph, rng, al, P.adj, adj_k, rmode, st);
        wave_lds_sync();
    }
    if (lane == 0)
        for (int k = 0; k < kStats; ++k)
            if (st[256 * k]) atomicAdd(P.stats + (k < 6 ? k : k + 4), st[256 * k]);
}

// Per (b,k,θ) cell, one wave: the cell's np outcome bytes in steps of 1 KB, 16 bytes per
// lane.  The count is the number of transmitted paths before the cell's first start-grain
// mismatch (the reference EXITs the particle loop there, RT:1673-1682); the paths after
// it are added to `skipped`.  The outcome buffer is padded by 16 bytes, so a lane's
// 16-byte block that starts inside it stays inside it.
__global__ void __launch_bounds__(256) reduce_grid_kernel(const uint8_t *__restrict__ outcome,
                                                          int64_t ncell, int64_t np, int64_t nb,
                                                          int64_t nk, int32_t nangle,
                                                          long long *counts,
                                                          unsigned long long *skipped)
{
    const int lane = threadIdx.x & 63;
    const int64_t cell = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (cell >= ncell) return;                       // whole waves
    const int64_t base = cell * np, end = base + np;
    long long cnt = 0;
    int64_t first = end;                             // first mismatch (absolute position)
    for (int64_t c = base & ~(int64_t)15; c < end; c += 1024) {
        const int64_t q = c + 16 * lane;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (q < end) v = *(const uint4 *)(outcome + q);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        int64_t lf = end;
        for (int j = 15; j >= 0; --j) {
            const unsigned byte = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
            if (q + j >= base && q + j < end && (byte & OUT_MISMATCH)) lf = q + j;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const int64_t t = __shfl_xor(lf, off, 64);
            lf = t < lf ? t : lf;
        }
        long long lc = 0;
        for (int j = 0; j < 16; ++j) {
            const unsigned byte = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
            lc += (q + j >= base && q + j < lf && (byte & OUT_TRANSMITTED)) ? 1 : 0;
        }
        for (int off = 32; off > 0; off >>= 1) lc += __shfl_xor(lc, off, 64);
        cnt += lc;
        if (lf < end) {
            first = lf;
            break;
        }
    }
    if (lane == 0) {
        // cell = ((ik-1)*nb + (ib-1))*(nangle-1) + (ia-1)  ->  transmit(ib,ik,ia)
        const int64_t na = nangle - 1;
        const int64_t ia = cell % na, rest = cell / na;
        const int64_t ib = rest % nb, ik = rest / nb;
        counts[ib + nb * (ik + nk * ia)] = cnt;
        if (first < end) atomicAdd(skipped, (unsigned long long)(end - first - 1));
    }
}

// Single paths with full records (parity checks, trajectory logs), one lane per path:
// path first + i * stride for i < n.  With pts, the lane also stores the positions the
// reference writes to trace.out: the start position (RT:1663, before the start-grain
// check) and the position at the end of every event-loop iteration (RT:1818-1822; a path
// stopped at the top of an iteration, as lost or at the event limit, writes nothing
// more); the first max_pts of them, at pts[(i * max_pts + k) * 3], and their number
// (all of them) at npts[i].
template <int kMode, int kPrec>
__global__ void __launch_bounds__(64) replay_kernel(TraceParams P, int64_t first, int64_t n,
                                                    int64_t stride, pt_path_record *rec,
                                                    double *pts, int64_t max_pts, int64_t *npts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    __shared__ unsigned long long unused_stats[8][256];      // event statistics, not reported
    OpCount oc = {0ull, 0ull, 0ull, 0ull, &unused_stats[0][threadIdx.x]};
    Rng rng;
    rng.oc = &oc;
    rng.seed = P.seed;
    Path ph;
    AdjLane al = {nullptr, 0, nullptr, 0, 0.0, nullptr, -1, 0.0, 0.0, 0.0};
    int mode = 0;
    int64_t cnt = 0;
    const bool started = begin_path<kMode, kPrec>(P, first + i * stride, rng, ph, &mode);
    if (pts) {
        double *q = pts + (i * max_pts) * 3;
        q[0] = ph.pos[0]; q[1] = ph.pos[1]; q[2] = ph.pos[2];
    }
    cnt = 1;
    if (started) {
        while (ph.end_side == 0) {
            const int64_t ev = ph.events;
            const bool done = step_path<kPrec, false>(P.g, P.m.max_omega, P.event_limit, rng, ph, al);
            if (ph.events != ev) {
                if (pts && cnt < max_pts) {
                    double *q = pts + (i * max_pts + cnt) * 3;
                    q[0] = ph.pos[0]; q[1] = ph.pos[1]; q[2] = ph.pos[2];
                }
                ++cnt;
            }
            if (done) break;
        }
    }
    if (npts) npts[i] = cnt;
    pt_path_record r;
    r.end_side = ph.end_side;
    r.flags = ph.flags;
    r.events = ph.events;
    r.draws = rng.draw;
    r.which_grain = ph.which;
    r.start_grain = ph.start_grain;
    for (int k = 0; k < 3; ++k) { r.pos[k] = ph.pos[k]; r.dir[k] = ph.dir[k]; }
    rec[i] = r;
}

// Workgroups of the persistent kernel that one tail workgroup displaces on a CU holding
// per_cu of them: per SIMD one wave of each workgroup (256 threads: four waves), 512 VGPRs
// per SIMD lane in granules of 8, lds_cu bytes of LDS per CU.  0 when the attributes are
// not available.
static int live_displace(const void *main_fn, size_t main_dyn_lds, const void *tail_fn, int per_cu,
                         size_t lds_cu)
{
    hipFuncAttributes am, at;
    if (!main_fn || !tail_fn || per_cu <= 0 || hipFuncGetAttributes(&am, main_fn) != hipSuccess ||
        hipFuncGetAttributes(&at, tail_fn) != hipSuccess)
        return 0;
    const int vm = (am.numRegs + 7) / 8 * 8, vt = (at.numRegs + 7) / 8 * 8;
    const size_t lm = am.sharedSizeBytes + main_dyn_lds, lt = at.sharedSizeBytes;
    for (int k = per_cu; k >= 0; --k)
        if (k * vm + vt <= 512 && k * lm + lt <= lds_cu) return per_cu - k;
    return 0;
}

hipError_t launch_trace(const TraceParams &p, int mode, int blocks, int threads, bool count,
                        size_t lds, hipStream_t s, const LiveTail *live)
{
    const bool hand = tail_mode(mode, p.precision) && !count && p.tail_cap > 0;
    if (hand && threads != 256) return hipErrorInvalidValue;   // four waves per workgroup
    using TailFn = void (*)(TraceParams);
    TailFn tail_fn = hand ? reinterpret_cast<TailFn>(const_cast<void *>(tail_kernel_fn(mode, p.precision)))
                          : nullptr;
    // the live tail kernel first, on its own stream once the launch's stream has reached
    // this point (its counters are cleared); the persistent kernel leaves it the room of
    // as many of its own workgroups per tail workgroup as share a CU's VGPRs and LDS with
    // it.  Launched before the persistent kernel is resident, it gets that room; launched
    // after, it waits for the persistent kernel's workgroups to end; if the persistent
    // kernel cannot become resident beside it, it steps aside after live_ticks / 200
    // (the persistent kernel waits for nothing).
    const int displace = hand && live && live->blocks > 0
                             ? live_displace(trace_kernel_fn(mode, p.precision), lds, tail_kernel_fn(mode, p.precision),
                                             blocks / (live->ncu > 0 ? live->ncu : 1), live->lds_cu)
                             : 0;
    const bool side = displace > 0 && blocks - displace * live->blocks >= blocks / 2;
    int main_blocks = blocks;
    hipError_t e = hipSuccess;
    if (side) {
        main_blocks = blocks - displace * live->blocks;
        TraceParams q = p;
        q.tail_live = 1;
        q.main_waves = (int64_t)main_blocks * (threads / 64);
        q.live_ticks = live->ticks;
        if ((e = hipEventRecord(live->fork, s))) return e;
        if ((e = hipStreamWaitEvent(live->side, live->fork, 0))) return e;
        hipLaunchKernelGGL(tail_fn, dim3(live->blocks), dim3(threads), 0, live->side, q);
        if ((e = hipGetLastError())) return e;
    }
    // the pre-pass of the queue's first pass (the live tail kernel takes its CUs meanwhile:
    // once the persistent kernel is resident, no CU has room for a tail workgroup)
    if (p.first_ids) {
        TraceParams q = p;
        q.live_gate = side ? live->blocks : 0;
        q.live_ticks = side ? live->ticks : 0;
        hipLaunchKernelGGL(first_scan_kernel, dim3(blocks), dim3(256), 0, s, q);
        if ((e = hipGetLastError())) return e;
    }
    const dim3 gr(main_blocks), bl(threads);
#define PT_LAUNCH(M, C, R) hipLaunchKernelGGL((trace_kernel<M, C, R>), gr, bl, lds, s, p)
#define PT_LAUNCH_C(M, R) do { if (count) PT_LAUNCH(M, true, R); else PT_LAUNCH(M, false, R); } while (0)
#define PT_LAUNCH_R(M)                                                    \
    do {                                                                  \
        if (p.precision == kPrecFast) PT_LAUNCH_C(M, kPrecFast);          \
        else if (p.precision == kPrecFilter) PT_LAUNCH_C(M, kPrecFilter); \
        else PT_LAUNCH_C(M, kPrecFP64);                                   \
    } while (0)
    if (mode == TRACE_GRID) PT_LAUNCH_R(TRACE_GRID);
    else if (mode == TRACE_GRID_TAIL && p.precision == kPrecFast) PT_LAUNCH_C(TRACE_GRID_TAIL, kPrecFast);
    else if (mode == TRACE_GRID_TAIL && p.precision == kPrecFilter) PT_LAUNCH_C(TRACE_GRID_TAIL, kPrecFilter);
    else if (mode == TRACE_GRID_TAIL) PT_LAUNCH_C(TRACE_GRID, kPrecFP64);
    else if (mode == TRACE_ALIAS) PT_LAUNCH_R(TRACE_ALIAS);
    else if (mode == TRACE_ADJOINT) PT_LAUNCH_R(TRACE_ADJOINT);
    else if (mode == TRACE_GRAINS) PT_LAUNCH_R(TRACE_GRAINS);
    else PT_LAUNCH_R(TRACE_BATCH);
#undef PT_LAUNCH_R
#undef PT_LAUNCH_C
#undef PT_LAUNCH
    e = hipGetLastError();
    if (e == hipSuccess && hand) {
        TraceParams q = p;
        q.tail_live = 0;
        hipLaunchKernelGGL(tail_fn, dim3(blocks), bl, 0, s, q);
        e = hipGetLastError();
    }
    if (side) {
        // the launch's stream waits for the live tail kernel too
        const hipError_t f = hipEventRecord(live->join, live->side);
        const hipError_t w = f ? f : hipStreamWaitEvent(s, live->join, 0);
        if (e == hipSuccess) e = w;
    }
    return e;
}

const void *tail_kernel_fn(int mode, int precision)
{
#define PT_TAIL(M) (precision == kPrecFast ? (const void *)tail_kernel<M, kPrecFast> \
                                           : (const void *)tail_kernel<M, kPrecFilter>)
    if (!tail_mode(mode, precision)) return nullptr;
    if (mode == TRACE_GRID_TAIL) return PT_TAIL(TRACE_GRID_TAIL);
    if (mode == TRACE_ALIAS) return PT_TAIL(TRACE_ALIAS);
    if (mode == TRACE_ADJOINT) return PT_TAIL(TRACE_ADJOINT);
    if (mode == TRACE_GRAINS) return PT_TAIL(TRACE_GRAINS);
    return PT_TAIL(TRACE_BATCH);
#undef PT_TAIL
}

hipError_t trace_occupancy(int mode, int precision, int threads, size_t lds, int *blocks_per_cu)
{
#define PT_OCC(M, R) hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, trace_kernel<M, false, R>, threads, lds)
#define PT_OCC_R(M)                                                                  \
    (precision == kPrecFast ? PT_OCC(M, kPrecFast)                                   \
                            : precision == kPrecFilter ? PT_OCC(M, kPrecFilter) : PT_OCC(M, kPrecFP64))
    if (mode == TRACE_GRID || (mode == TRACE_GRID_TAIL && precision == kPrecFP64))
        return PT_OCC_R(TRACE_GRID);
    if (mode == TRACE_GRID_TAIL)
        return precision == kPrecFast ? PT_OCC(TRACE_GRID_TAIL, kPrecFast) : PT_OCC(TRACE_GRID_TAIL, kPrecFilter);
    if (mode == TRACE_ALIAS) return PT_OCC_R(TRACE_ALIAS);
    if (mode == TRACE_ADJOINT) return PT_OCC_R(TRACE_ADJOINT);
    if (mode == TRACE_GRAINS) return PT_OCC_R(TRACE_GRAINS);
    return PT_OCC_R(TRACE_BATCH);
#undef PT_OCC_R
#undef PT_OCC
}

const void *trace_kernel_fn(int mode, int precision)
{
#define PT_FN(M) (precision == kPrecFast ? (const void *)trace_kernel<M, false, kPrecFast>            \
                  : precision == kPrecFilter ? (const void *)trace_kernel<M, false, kPrecFilter>      \
                                             : (const void *)trace_kernel<M, false, kPrecFP64>)
    if (mode == TRACE_GRID || (mode == TRACE_GRID_TAIL && precision == kPrecFP64)) return PT_FN(TRACE_GRID);
    if (mode == TRACE_GRID_TAIL)
        return precision == kPrecFast ? (const void *)trace_kernel<TRACE_GRID_TAIL, false, kPrecFast>
                                      : (const void *)trace_kernel<TRACE_GRID_TAIL, false, kPrecFilter>;
    if (mode == TRACE_ALIAS) return PT_FN(TRACE_ALIAS);
    if (mode == TRACE_ADJOINT) return PT_FN(TRACE_ADJOINT);
    if (mode == TRACE_GRAINS) return PT_FN(TRACE_GRAINS);
    return PT_FN(TRACE_BATCH);
#undef PT_FN
}

// Zero fill of n 64-bit words (used instead of memset nodes in captured graphs).
__global__ void fill_zero_kernel(unsigned long long *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

const void *fill_kernel_fn() { return (const void *)fill_zero_kernel; }

hipError_t launch_fill(unsigned long long *p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_reduce_grid(const uint8_t *outcome, int64_t ncell, int64_t np, int64_t nb,
                              int64_t nk, int32_t nangle, long long *counts,
                              unsigned long long *skipped, hipStream_t s)
{
    if (ncell == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_grid_kernel, dim3((unsigned)((ncell + 3) / 4)), dim3(256), 0, s,
                       outcome, ncell, np, nb, nk, nangle, counts, skipped);
    return hipGetLastError();
}

hipError_t launch_replay(const TraceParams &p, int mode, int64_t first, int64_t n, int64_t stride,
                         pt_path_record *rec, double *pts, int64_t max_pts, int64_t *npts,
                         hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const dim3 gr((unsigned)((n + 63) / 64)), bl(64);
#define PT_REPLAY(M, R) hipLaunchKernelGGL((replay_kernel<M, R>), gr, bl, 0, s, p, first, n, stride, \
                                           rec, pts, max_pts, npts)
#define PT_REPLAY_R(M)                                                   \
    do {                                                                 \
        if (p.precision == kPrecFast) PT_REPLAY(M, kPrecFast);           \
        else if (p.precision == kPrecFilter) PT_REPLAY(M, kPrecFilter);  \
        else PT_REPLAY(M, kPrecFP64);                                    \
    } while (0)
    if (mode == TRACE_GRID) PT_REPLAY_R(TRACE_GRID);
    else PT_REPLAY_R(TRACE_ALIAS);
#undef PT_REPLAY_R
#undef PT_REPLAY
    return hipGetLastError();
}

}  // namespace pt
