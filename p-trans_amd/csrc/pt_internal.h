// pt_internal.h -- device data layout shared by the kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ptrans.h"

namespace pt {

// One surface as the trace kernels read it (the fields of Surface_type,
// src/RTMCphon.f90:94-111, that calcTransmit touches).  112 bytes, 16-byte aligned.
struct alignas(16) DSurf {
    double n[3];        // polygon normal (not renormalised, RT:1459)
    double c[3];        // centre (sphere/cylinder) or centroid (polygon)
    double r;           // radius
    double p;           // specularity
    double tparam;      // transmissivity law parameter (t or gamma)
    int32_t kind;       // PT_SHAPE_*
    int32_t nv;         // polygon vertex count
    int32_t voff;       // first vertex in the vertex pool
    int32_t gout;       // grain_outside
    int32_t tkind;      // PT_TRANS_*
    int32_t gin;        // grain_inside
    int32_t pad[2];
};
static_assert(sizeof(DSurf) == 112, "DSurf layout");

// One polygon candidate in scan order (the grain's candidate list): the plane data the
// filter and the exact plane distance read, so one 64-byte load per candidate needs no
// index chain.  idx: the global surface index.
struct alignas(16) CandPlane {
    double n[3];        // the surface's normal (as in DSurf)
    double v1[3];       // its first vertex
    int32_t idx;
    int32_t pad[3];
};
static_assert(sizeof(CandPlane) == 64, "CandPlane layout");

// Geometry in HBM.  Candidate lists hold, per grain, the global indices of its
// internal surfaces in reference order (the scan of RT:2012-2028 skips the others).
struct DevGeom {
    const DSurf *surf;
    const double *tang;        // 9 doubles per surface, column-major (tangent(:,j))
    const double *vtx;         // xyz triples
    const int32_t *cand_off;   // [ngrain_total+1]
    const int32_t *cand_idx;
    const int32_t *inc_off;    // incident surfaces per grain (findIncidentCollisionGrain)
    const int32_t *inc_idx;
    const int32_t *grain_incident;
    const double *gen;         // generatrix xyz per grain
    const int32_t *surf_base;  // global index of each grain's first surface
    const CandPlane *cand_plane;  // [cand_off[ngrain_total]] when lane_scan
    int32_t lane_scan;         // every grain candidate is a polygon and the box is aligned:
                               // the accelerated scan runs each lane's own candidate list
    double box[3];
    double scope[2];
    int32_t ngrain;
    int32_t kind_grain;
    int32_t igrain_fill;
    int32_t nmaterial;
    int32_t ngrain_total;
    int32_t nsurface;
    int32_t box_aligned;       // grain 0 is the six axis-aligned walls of setupBoxGrain
    int32_t periodic;          // some side walls are PT_TRANS_PERIODIC (box_aligned too)
};

struct DevModes {
    const double *omega;       // (nb, nk) column-major
    const double *mfp;         // (nb, nk) mean free paths [nm] (alias mode)
    const double *prob;        // Vose table (alias mode)
    const int32_t *alias;
    const double *scale;       // per-branch scattering-rate scale or nullptr
    int32_t nb, nk;
    double max_omega;
};

// Per-path outcome bytes of the grid trace.
enum : uint8_t {
    OUT_TRANSMITTED = 1,       // EndSide /= StartSide (RT:1833)
    OUT_MISMATCH = 2,          // start-grain check failed (RT:1673-1682)
};

// TRACE_ADJOINT: the alias-mode trace plus likelihood-ratio scores of each path;
// TRACE_BATCH: the adjoint-mode trace over a batch of mesh structures (BVH scans);
// TRACE_GRID_TAIL: the grid trace of a diffusive sampling mean free path (one that gets
// safety balls, below kBallMfpFrac of the smallest box side), with the few-path service at
// the end of each wave.  The ballistic grid trace keeps the leaner TRACE_GRID kernel.
// TRACE_GRAINS: the alias-mode trace plus per-grain flux tallies (no likelihood ratio).
enum { TRACE_GRID = 0, TRACE_ALIAS = 1, TRACE_ADJOINT = 2, TRACE_BATCH = 3, TRACE_GRID_TAIL = 4,
       TRACE_GRAINS = 5 };

constexpr bool is_grid_mode(int mode) { return mode == TRACE_GRID || mode == TRACE_GRID_TAIL; }

// Safety balls (accelerated precisions) are set up for paths whose mean free path is below
// this fraction of the smallest box side.
constexpr double kBallMfpFrac = 0.05;

// Binary BVH node holding both children's bounds (outward padded, FP32), so one load
// decides which children a ray or a ball query enters and in which order, and, per
// child, FP32 bounds of its prims' normals (outward rounded; [-1, 1] for curved prims): a
// child whose every polygon certainly faces away from the ray (the reference's cosang <= 0,
// a miss) is not entered.  128 bytes: one cache line.
struct BvhNode {
    float lo[2][3];
    float hi[2][3];
    int32_t child[2];   // >= 0: internal node index; < 0: leaf, prims from -(child + 1)
    int32_t count[2];   // leaf prim count (0 for an internal child)
    float nlo[2][3];    // per child: componentwise bounds of the prims' normals
    float nhi[2][3];
    int32_t pad[4];
};
static_assert(sizeof(BvhNode) == 128, "BvhNode layout");
constexpr int kBvhStack = 16;          // traversal stack per lane (LDS); tree depth < 16
// per-structure sums: n, sum x, sum x^2, sum (x - b) score, sum ((x - b) score)^2
constexpr int kBatchSums = 5;
// Batch work queues, one per XCD: each covers a contiguous range of structures, so the
// L2 of an XCD serves the BVHs and facets of the few structures its waves trace at a
// time; a wave whose own queue is empty takes chunks from the others.
constexpr int kXcdQueues = 8;

// A batch of structures merged into one mesh geometry (shared box grain, disjoint grain
// sets): per structure its path range, start locator and tallies; per grain the BVH of
// its candidate surfaces.
struct DevBatch {
    int32_t nstruct;
    int32_t grid;                      // start locator: grid x grid cells over y, z
    const int64_t *path_off;           // [nstruct + 1]
    const int32_t *grain_root;         // [ngrain_total]: BVH root of the grain, -1 none
    const BvhNode *nodes;
    const int32_t *prims;              // global surface indices, leaf order
    const int32_t *loc_off;            // [nstruct * grid * grid + 1]
    const int32_t *loc_idx;            // incident facets overlapping each cell
    const double *baseline;            // [nstruct]
    double *sums;                      // [nstruct * kBatchSums]
    unsigned int *done;                // [nstruct] paths finished
    unsigned long long *finish;        // [nstruct] wall clock when the last path ended
    unsigned long long *start;         // wall clock at kernel start (one word)
    const int64_t *qoff;               // [kXcdQueues + 1]: path range of each queue
    unsigned long long *qctr;          // [kXcdQueues]: paths taken from each queue
};

// Largest number of specularity groups of one adjoint launch (each lane keeps one FP64
// score per group in LDS).
constexpr int kAdjMaxGroups = 32;

struct TraceParams {
    DevGeom g;
    DevModes m;
    const double *cos_t;       // grid: per angle index
    const double *sin_t;
    uint64_t seed;
    double mfp;                // grid: the sampling mean free path (1e10 ballistic)
    int64_t np;                // grid: particles per cell
    int64_t total;             // paths of the launch
    int32_t nangle;            // grid
    int32_t chunk;             // paths a wave takes from the queue at once
    int32_t refill_min;        // idle lanes that trigger a refill (64: only when all idle)
    int64_t event_limit;
    int32_t precision;         // PT_PREC_FP64 or PT_PREC_FP32_RETEST
    unsigned long long *queue; // next path id
    unsigned long long *stats; // paths, events, draws, lost, mismatch, limit, (skipped),
                               // fp64 ops, philox blocks (counting kernel only),
                               // FP64 retests, full-scan fallbacks
    uint8_t *outcome;          // grid: [total]
    unsigned long long *mode_paths;   // alias: [nb*nk]
    unsigned long long *mode_trans;   // alias: [nb*nk]
    // adjoint: per batch (path id mod nbatch) the sums [n, sum x, sum x^2,
    // sum (x - baseline) * score_k for k < ngroup + nb], x = T / mfp_m
    const int32_t *surf_group; // [nsurface]: specularity group of each surface, or -1
    int32_t ngroup;
    int32_t nbatch;
    double baseline;
    double *adj;               // [nbatch * (3 + ngroup + nb)]
    // per-grain tallies (adjoint mode, grain_tally != nullptr): per grain g the sums of
    // dx / mfp_m and l / mfp_m over the flights inside g, [2 * g], [2 * g + 1]
    double *grain_tally;
    int32_t ngt;               // grains tallied (ngrain_total) or 0
    DevBatch bt;               // TRACE_BATCH
};

__host__ __device__ inline int adj_stride(const TraceParams &p) { return 3 + p.ngroup + p.m.nb; }
// Static LDS of trace_kernel: eight per-lane event counters of a 256-lane workgroup.
constexpr size_t kTraceStaticLds = 8 * 256 * sizeof(unsigned long long);
// Largest event limit: a path draws at most 7 numbers before its first event and 6 per
// event (specular decision, two direction draws, the transmission draw, the nmaterial > 1
// draw of RT:2399 and a new free path), and its draw counter is 32 bits wide.
constexpr int64_t kMaxEventLimit = (((int64_t)1 << 32) - 8) / 6;
// Dynamic LDS of trace_kernel (the static kTraceStaticLds comes on top)
inline size_t adj_lds_bytes(const TraceParams &p, int threads)
{
    return ((size_t)p.ngroup * threads + (size_t)p.nbatch * adj_stride(p) + 2 * (size_t)p.ngt) *
               sizeof(double) +
           (p.bt.nstruct ? (size_t)kBvhStack * threads * sizeof(int32_t) : 0);
}

// kernels' launch wrappers (pt_trace.hip)
hipError_t launch_trace(const TraceParams &p, int mode, int blocks, int threads, bool count,
                        size_t lds, hipStream_t s);
hipError_t launch_reduce_grid(const uint8_t *outcome, int64_t ncell, int64_t np,
                              int64_t nb, int64_t nk, int32_t nangle,
                              long long *counts, unsigned long long *skipped, hipStream_t s);
hipError_t launch_replay(const TraceParams &p, int mode, int64_t first, int64_t n, int64_t stride,
                         pt_path_record *rec, double *pts, int64_t max_pts, int64_t *npts,
                         hipStream_t s);
// workgroups per CU the trace kernel of (mode, precision) can hold with `lds` bytes of
// dynamic LDS
hipError_t trace_occupancy(int mode, int precision, int threads, size_t lds, int *blocks_per_cu);
// the trace kernel of (mode, precision) and the zero-fill kernel, for graph nodes
const void *trace_kernel_fn(int mode, int precision);
const void *fill_kernel_fn();
hipError_t launch_fill(unsigned long long *p, size_t n, hipStream_t s);

}  // namespace pt
