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
    float nf[3];        // the normal rounded to FP32 (the filter's)
    int32_t idx;
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
    // lane_scan geometries: per grain g the planes of its scan order (the box grain's
    // candidates, then g's), unit normal and offset (n, n.v) in FP32, oriented so that the
    // grain's generatrix lies on the positive side: [region_off[g], region_off[g + 1]);
    // a grain whose generatrix lies on one of them gets none
    const float4 *region;
    const int32_t *region_off;
    float region_margin;       // distance to every plane that a region test requires (nm)
    // Voronoi geometries of polygon grains: the start grain of every point of a cell of a
    // start_grid x start_grid grid over the start face (y, z), or 0 where the cell needs
    // the full searches (findNearestIncidentGrain and the start-grain check)
    const int32_t *start_map;
    int32_t start_grid;
    double start_ih[2];        // inverse cell sides in y and z
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
    const double *inv_mfp;     // 1.0 / mfp, correctly rounded (alias mode)
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
// safety balls, below kBallMfpFrac of the smallest box side), with its long walks and the
// last few paths of each wave handed to the tail kernel.  The ballistic grid trace keeps the
// leaner TRACE_GRID kernel.
// TRACE_GRAINS: the alias-mode trace plus per-grain flux tallies (no likelihood ratio).
enum { TRACE_GRID = 0, TRACE_ALIAS = 1, TRACE_ADJOINT = 2, TRACE_BATCH = 3, TRACE_GRID_TAIL = 4,
       TRACE_GRAINS = 5 };

constexpr bool is_grid_mode(int mode) { return mode == TRACE_GRID || mode == TRACE_GRID_TAIL; }

// Safety balls (accelerated precisions) are set up, by the tail kernel only, for paths whose
// mean free path is below this fraction of the smallest box side.
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
// Radius of a tail walk's local candidate list (batch): this many mean free paths, at least
// kLocalMin nm (a quarter, then a sixteenth of it when more than 64 surfaces lie within)
constexpr float kLocalMfps = 12.0f;
constexpr float kLocalMin = 6.0f;

// Top levels of the grain BVHs of one structure staged in a workgroup's LDS.  A node
// reference r >= kLdsRef names LDS slot r - kLdsRef (grain i of the structure, heap
// position j of its top levels: slot i * K + j, K = 2^levels - 1); below it, a global
// node index.  Slots are 144 bytes apart, so that lanes reading the roots of different
// grains spread over the LDS banks.
constexpr int32_t kLdsRef = 1 << 30;
constexpr int kMaxTopLevels = 3;
// heap slots per grain in the device image the LDS copies are made from
constexpr int kTopImage = (1 << kMaxTopLevels) - 1;
// path flag (internal, never reported): the path reads the LDS image of its structure
constexpr int kPathTop = 1 << 30;
// a path bound to the LDS image leaves it after this many events (a long random walk)
constexpr int64_t kTopUnbindEvents = 2000;
struct TopSlot {
    BvhNode n;
    int32_t pad[4];
};
static_assert(sizeof(TopSlot) == 144, "TopSlot layout");
// Header of the LDS image: word = bound lanes (bits 0-15) | structure (bits 16-30; kTopNone:
// none) | being staged (bit 31); the structure's first global grain and its grains staged.
struct TopHdr {
    unsigned int word;
    int32_t g0, ng, pad;
};
constexpr unsigned int kTopNone = 0x7fffu;
constexpr unsigned int kTopBusy = 0x80000000u;
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
    // LDS staging of the top BVH levels (top_k = 0: off): per global grain its top_k
    // nodes in heap order with LDS child references, [ngrain_total * top_k]; the first
    // global grain of each structure, [nstruct + 1]; grains staged per structure
    const BvhNode *top;
    const int32_t *sgrain;
    int32_t top_k;
    int32_t top_cap;
};

// Largest number of specularity groups of one adjoint launch (each lane keeps one FP64
// score per group in LDS).
constexpr int kAdjMaxGroups = 32;

// A long random walk handed from the persistent trace kernel to the tail kernel: the
// whole state of the path between two events (the per-group likelihood-ratio scores go
// to a separate array, ngroup doubles per record).  192 bytes.
struct alignas(16) TailRec {
    double pos[3];
    double dir[3];
    double fp, mfp, omega;
    double spare;               // the Rng's second half-block, when draw is odd
    double len;                 // adjoint: the path's flight length so far
    double fdx, fl, fw;         // grain tallies: the open grain's sums and 1 / mfp_m
    int64_t pid, events;
    float bc[3], br;
    int32_t which, cur, flags;
    uint32_t epoch;             // the launch that wrote the record, stored last
    int32_t mode, sid, nscat, fgrain;     // sid: batch traces only
    uint32_t draw;
    int32_t pad[3];
};
static_assert(sizeof(TailRec) == 192, "TailRec layout");
// The counter block of a launch: queue at 0, statistics at 1-14 (pt_stats order, see
// fill_stats), then the tail records claimed in the general and the priority part of the
// buffer (either may count past its capacity: those paths stayed), the tail kernels' next
// record of each part and the waves of the persistent kernel that have ended.
constexpr int kCounterWords = 32;
constexpr int kCtrTailN = 16;
constexpr int kCtrTailP = 17;
constexpr int kCtrTailNext = 18;
constexpr int kCtrTailNextN = 19;
constexpr int kCtrMainDone = 20;
constexpr int kCtrFirstN = 21;       // paths the pre-pass found for the first pass
constexpr int kCtrMainStarted = 22;  // waves of the persistent kernel that have begun
constexpr int kCtrLiveStarted = 23;  // live tail workgroups that have begun
// Paths whose mean free path is below this fraction of the smallest box side take records
// in the priority part of the tail buffer, which the tail kernels serve first: their walks
// are the longest (~(L / mfp)^2 events), so they start at once.  The other paths are handed
// over only after kTailLongFactor times as many events: most of their walks end within a
// few thousand events, which the persistent kernel runs at a fraction of the tail kernel's
// cost per event.
constexpr double kTailPrioMfpFrac = 5.0e-3;
constexpr int64_t kTailLongFactor = 4;
// Alias-mode launches with a tail kernel start the paths of the last quarter of the ids
// whose mean free path is below this fraction of the smallest box side before the rest of
// that quarter (a pre-pass lists them; the persistent kernel's queue runs the ids of the
// first three quarters, the list, then the last quarter without the listed paths): the
// longest walks then reach the live tail kernel while the persistent kernel has at least a
// quarter of its work left, not in its last moments.
constexpr double kFirstMfpFrac = 2.0e-3;

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
    // long random walks (accelerated precisions, every mode but the ballistic grid): a path
    // past tail_events events, or one of the last few of a wave once the queue is empty, is
    // handed to the tail kernel, which runs it on a whole wave; at most tail_cap records
    // (0: none, paths stay in the persistent kernel)
    TailRec *tail;
    double *tail_score;        // [tail_cap * ngroup] (adjoint and batch modes)
    int64_t tail_cap;          // records in all; the first tail_cap_p are the priority part
    int64_t tail_cap_p;
    double tail_prio_mfp;      // paths below this mean free path take priority records
    int64_t tail_events;
    uint32_t tail_epoch;       // this launch's number, written last into each record
    // the tail kernel beside the persistent one (on another stream): it takes records while
    // they are handed over, until all main_waves waves have ended or live_ticks of the wall
    // clock (s_memrealtime) have passed, or after live_ticks / 200 if not every wave of the
    // persistent kernel has begun (launches serialised, as under a counter profiler, or
    // workgroups of the persistent kernel waiting for the room the live ones hold)
    int32_t tail_live;
    int32_t live_gate;         // the pre-pass waits for this many live workgroups (0: none)
    int64_t main_waves;
    int64_t live_ticks;
    // the queue's list of the paths from first_from on below first_mfp (at most first_cap;
    // a launch that finds more runs every path in id order)
    int64_t *first_ids;
    int64_t first_cap;
    int64_t first_from;
    double first_mfp;
};

// Modes and precisions whose long walks go to the tail kernel (the op-counting kernels
// keep them, so that every operation is counted in one place).
constexpr bool tail_mode(int mode, int precision)
{
    return mode != TRACE_GRID && precision != PT_PREC_FP64;
}

__host__ __device__ inline int adj_stride(const TraceParams &p) { return 3 + p.ngroup + p.m.nb; }
// Static LDS of trace_kernel: per-lane event counters of a 256-lane workgroup, eight of
// them (the batch kernel, which has no FP64 retests or scan fallbacks: six).
constexpr int trace_lane_stats(int mode) { return mode == TRACE_BATCH ? 6 : 8; }
constexpr size_t trace_static_lds(int mode)
{
    return (size_t)trace_lane_stats(mode) * 256 * sizeof(unsigned long long);
}
constexpr size_t kTraceStaticLds = trace_static_lds(TRACE_GRID);
// Largest event limit: a path draws at most 7 numbers before its first event and 6 per
// event (specular decision, two direction draws, the transmission draw, the nmaterial > 1
// draw of RT:2399 and a new free path), and its draw counter is 32 bits wide.
constexpr int64_t kMaxEventLimit = (((int64_t)1 << 32) - 8) / 6;
// Dynamic LDS of trace_kernel (the static kTraceStaticLds comes on top)
// Layout: [ngroup * threads scores | nbatch * adj_stride sums | 2 * ngt grain tallies]
// (doubles), then for batches, from a 16-byte boundary, [kBvhStack * threads stack
// entries | TopHdr | top_k * top_cap TopSlots].
__host__ __device__ inline size_t adj_lds_doubles(const TraceParams &p, int threads)
{
    return (size_t)p.ngroup * threads + (size_t)p.nbatch * adj_stride(p) + 2 * (size_t)p.ngt;
}
__host__ __device__ inline size_t batch_lds_offset(const TraceParams &p, int threads)
{
    return (adj_lds_doubles(p, threads) * sizeof(double) + 15) & ~(size_t)15;
}
inline size_t adj_lds_bytes(const TraceParams &p, int threads)
{
    if (!p.bt.nstruct) return adj_lds_doubles(p, threads) * sizeof(double);
    return batch_lds_offset(p, threads) + (size_t)kBvhStack * threads * sizeof(int32_t) +
           (p.bt.top_k ? sizeof(TopHdr) + (size_t)p.bt.top_k * p.bt.top_cap * sizeof(TopSlot) : 0);
}

// The tail kernel beside a persistent launch: `blocks` workgroups on stream `side`, forked
// from and joined back into the launch's stream by the two events.  The persistent kernel
// gets as many workgroups fewer as make room for them (from both kernels' VGPRs and LDS).
struct LiveTail {
    hipStream_t side;
    hipEvent_t fork, join;
    int blocks;
    int64_t ticks;             // the live tail kernel's wall-clock limit (s_memrealtime)
    int ncu;                   // CUs of the device and LDS bytes per CU
    size_t lds_cu;
};

// kernels' launch wrappers (pt_trace.hip)
hipError_t launch_trace(const TraceParams &p, int mode, int blocks, int threads, bool count,
                        size_t lds, hipStream_t s, const LiveTail *live = nullptr);
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
// the tail kernel of (mode, precision) (nullptr: that trace hands nothing over)
const void *tail_kernel_fn(int mode, int precision);
const void *fill_kernel_fn();
hipError_t launch_fill(unsigned long long *p, size_t n, hipStream_t s);

}  // namespace pt
