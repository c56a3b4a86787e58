/*
 * ptrans.h -- C-ABI of libptrans, the gfx950 phonon Monte Carlo trace.
 *
 * Drop-in boundary.  In P-TRANS the trace is the Fortran module procedure
 *   subroutine calcTransmit(trans, cbranch, ckpoint, angle)   (src/RTMCphon.f90:1608-2414)
 * driven by
 *   subroutine TransmissionFunction                            (src/RTMCphon.f90:1544-1604)
 * whose OpenMP loop (RT:1573-1582) fills transmit(nbranch,nkpoint,nangle) and whose
 * MPI_Allreduce (RT:1584-1590) sums it over ranks.  libptrans replaces the body of
 * TransmissionFunction with one call, pt_trace_grid(), that runs every (k, b, angle)
 * cell x Np particles on one GPU, one phonon path per lane.  The Fortran host keeps
 * loadMCP (RT:252-1251), the .anphon reader (RT:661-733), effectiveMFP (RT:2418-2451)
 * and fileoutput (RT:2601-2881); it hands over the geometry it built (vgrain, RT:122)
 * flattened into the structures below, and the bulk phonon table (anphon, RT:174).
 *
 * Conventions
 *  - Every function returns PT_OK (0) or a negative PT_E_* code; pt_last_error()
 *    returns a message for the calling thread.  Nothing is printed.
 *  - Host arrays passed in are copied at create time; handles own device memory.
 *    Outputs are caller-allocated host arrays.
 *  - Calls on one context are not re-entrant; each blocking call synchronises the
 *    context's stream before it returns.
 *  - Arrays indexed (branch, kpoint) are Fortran column-major: element (ib, ik),
 *    1-based in the reference, is at [(ib-1) + nbranch*(ik-1)].
 *  - Random numbers: Philox4x32-10, one stream per phonon path, keyed by the context
 *    seed; results do not depend on the GPU, the launch shape or the order of paths.
 */
#ifndef PTRANS_H
#define PTRANS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 1

/* status codes */
#define PT_OK             0
#define PT_E_INVALID     -1   /* bad argument */
#define PT_E_HIP         -2   /* HIP runtime error */
#define PT_E_NOMEM       -3   /* device or host allocation failed */
#define PT_E_GEOMETRY    -4   /* inconsistent geometry descriptor */
#define PT_E_UNSUPPORTED -5   /* feature not available in this build */
#define PT_E_NODEVICE    -6   /* no usable GPU */

/* surface shapes, src/RTMCphon.f90:72-74 (the cylinder axis is z, RT:3051-3085) */
#define PT_SHAPE_POLYGON  1
#define PT_SHAPE_SPHERE   2
#define PT_SHAPE_CYLINDER 3

/* transmissivity laws, src/RTMCphon.f90:77-79 and 1402-1429 */
#define PT_TRANS_VALUE   1   /* t = trans_param                                    */
#define PT_TRANS_GAMMA   2   /* t = 1/(gamma*omega/omega_max + 1), gamma=trans_param */
#define PT_TRANS_TBC     3   /* as GAMMA; gamma was solved from a TBC by the host    */
/* extensions (no counterpart in the reference, whose laws are the three above):      */
#define PT_TRANS_LORENTZ 4   /* t = 1/(1 + (omega/omega0)^2), omega0 = trans_param
                                (rad/s), decided by u <= t as the reference's laws    */
#define PT_TRANS_PERIODIC 5  /* box side walls only (surfaces 2..5 of grain 0, in
                                pairs): a phonon reaching the wall continues from the
                                opposite wall with the same direction and no draw;
                                inclusions must not cross a periodic wall.  The
                                reference has no periodic boundary (its box walls
                                reflect with CellSpecularity, RT:1255-1331).          */

/* grain kinds, src/RTMCphon.f90:131-132 */
#define PT_GRAIN_VORONOI 1
#define PT_GRAIN_BRICKS  2
/* extension: grains that tile the box with triangulated facets (a jittered lattice);
   the start grain is the grain of the incident facet containing the start point
   (the rule of findIncidentCollisionGrain, RT:2131-2163, without a fill grain) and
   there is no isInsideGrain check (RT:1673-1682), which needs convex grains */
#define PT_GRAIN_MESH    3

/* arithmetic of the trace kernel */
#define PT_PREC_FP64 0           /* every step in FP64 (the reference's arithmetic)      */
#define PT_PREC_FP32_RETEST 1    /* FP32 intersection filter (bounded errors; grazing,
                                    edge, tangent and near-tie cases re-tested in FP64),
                                    the event's hit distance recomputed in FP64, and
                                    direction sampling and reflection in FP32; positions,
                                    free paths and tallies in FP64.  Statistically
                                    equivalent to FP64, not bit-identical.              */
#define PT_PREC_FP32_FILTER 2    /* the same FP32 filter with every other step in FP64:
                                    bit-identical to PT_PREC_FP64                       */

/*
 * One surface record: the fields of Surface_type (src/RTMCphon.f90:94-111) that the
 * trace reads.  Polygon vertices live in pt_geom_desc.vertices; the closing vertex
 * (vertex(:,nvertex+1) = vertex(:,1), RT:1093) is implicit.  normal, tangent and the
 * polygon centre are the values the host's setupSurfaceGeometry produced (RT:1433-1491);
 * pt_surface_setup_polygon() computes them the same way for hosts that need it.
 */
typedef struct pt_surface {
    int32_t kind;            /* PT_SHAPE_*                                      */
    int32_t nvertex;         /* polygon vertex count (>= 3); 0 otherwise        */
    int64_t vertex_offset;   /* first vertex, in units of 3 doubles             */
    int32_t internal;        /* surface%internal (only internal ones are hit)   */
    int32_t incident;        /* surface%incident                                */
    int32_t grain_inside;    /* surface%grain_inside                            */
    int32_t grain_outside;   /* surface%grain_outside                           */
    int32_t trans_kind;      /* PT_TRANS_*                                      */
    int32_t reserved;
    double  trans_param;     /* VALUE: t;  GAMMA/TBC: gamma                     */
    double  specularity;     /* surface%specularity                             */
    double  center[3];       /* surface%center                                  */
    double  radius;          /* surface%radius                                  */
    double  normal[3];       /* surface%normal (not renormalised, RT:1459)      */
    double  tangent[9];      /* surface%tangent(3,3), column-major              */
} pt_surface;

/*
 * The geometry: box, grains, surfaces.  Grain g (0-based, g = 0 is the box grain
 * igrain_box, RT:123) owns surfaces [grain_offset[g], grain_offset[g+1]) in the order
 * of vgrain(g)%vsurface; the collision scan visits grain 0 then the current grain in
 * that order and keeps the first strictly nearest hit (RT:2012-2028).
 */
typedef struct pt_geom_desc {
    double  box[3];              /* param%boxx, boxy, boxz (nm)                        */
    double  scope[2];            /* param%scopey, scopez                               */
    int32_t kind_grain;          /* PT_GRAIN_*                                         */
    int32_t ngrain;              /* ngrain (grains 1..ngrain; bricks add fill grain)   */
    int32_t igrain_fill;         /* ngrain+1 for bricks, -1 for voronoi                */
    int32_t nmaterial;           /* param%nmaterial                                    */
    int32_t ngrain_total;        /* entries of grain_offset minus one                  */
    int32_t nsurface;            /* entries of surfaces                                */
    int64_t nvertex;             /* entries of vertices / 3                            */
    const int32_t    *grain_offset;   /* [ngrain_total+1]                              */
    const int32_t    *grain_incident; /* [ngrain_total]  grain%incident                */
    const double     *generatrix;     /* [ngrain_total*3] grain%generatrix             */
    const pt_surface *surfaces;       /* [nsurface]                                    */
    const double     *vertices;       /* [nvertex*3]                                   */
} pt_geom_desc;

/* Bulk phonon table (Anphon_type, src/RTMCphon.f90:161-174), column-major (nb, nk).
 * The grid trace reads only omega and max_omega.  The alias-mode trace (an extension)
 * also needs vel, mfp and weight, and samples mode m with probability proportional to
 * C_m*v_m*mfp_m*weight_m, C the mode heat capacity at `temperature` (RT:2696-2701). */
typedef struct pt_modes_desc {
    int32_t nbranch;
    int32_t nkpoint;
    const double *omega;     /* rad/s, after the cm^-1 conversion of RT:727          */
    const double *vel;       /* m/s (alias mode; may be NULL for the grid trace)    */
    const double *mfp;       /* bulk mean free path in nm (alias mode; may be NULL) */
    const double *weight;    /* multiplicity (alias mode; may be NULL)              */
    double max_omega;        /* anphon%maxOmega (RT:728)                            */
    double temperature;      /* K, for the heat capacity of the alias weights       */
} pt_modes_desc;

/* Counters filled by the trace calls. */
typedef struct pt_stats {
    int64_t paths;           /* paths traced                                        */
    int64_t events;          /* event-loop iterations (RT:1693-1824), all paths      */
    int64_t draws;           /* random numbers consumed                             */
    int64_t lost;            /* paths that left the box (RT:1696-1701)               */
    int64_t start_mismatch;  /* start-grain check failures (RT:1673-1682)            */
    int64_t skipped;         /* particles not traced because of such a failure       */
    double  kernel_ms;       /* device time of the trace kernel (HIP events)         */
    double  reduce_ms;       /* device time of the per-cell tally reduction          */
    int64_t fp64_ops;        /* algorithmic FP64 operations (op counting on only)    */
    int64_t philox_blocks;   /* Philox4x32-10 blocks evaluated (op counting on only) */
    int64_t retests;         /* PT_PREC_FP32_RETEST: candidates settled by FP64       */
    int64_t fallbacks;       /* PT_PREC_FP32_RETEST: events decided by the FP64 scan  */
    int64_t bvh_nodes;       /* batch trace: BVH nodes visited (op counting on only)  */
    int64_t bvh_tests;       /* batch trace: leaf surfaces tested (op counting on)    */
    int64_t event_limited;   /* paths stopped by pt_ctx_set_event_limit's limit,
                                counted as not transmitted                           */
} pt_stats;

/* Per-path record of pt_replay(). */
typedef struct pt_path_record {
    int32_t end_side;        /* EndSide: 1 = x- (reflected), 2 = x+, 0 = left the box */
    int32_t flags;           /* PT_PATH_* bits                                       */
    int64_t events;          /* event-loop iterations (collision-1, RT:1659, 1705)   */
    int64_t draws;           /* random numbers consumed by the path                  */
    int32_t which_grain;     /* whichV at the end                                    */
    int32_t start_grain;     /* whichV after the start-grain search                  */
    double  pos[3];          /* final position (nm)                                  */
    double  dir[3];          /* final direction                                      */
} pt_path_record;

#define PT_PATH_LOST           1   /* left the box: counted as transmitted (RT:1833) */
#define PT_PATH_START_MISMATCH 2   /* start-grain check failed (RT:1673-1682)         */
#define PT_PATH_EVENT_LIMIT    4   /* stopped at the event safety limit               */

typedef struct pt_ctx   pt_ctx;
typedef struct pt_geom  pt_geom;
typedef struct pt_modes pt_modes;

/* ---- library ---------------------------------------------------------------- */

/* ABI version (PT_ABI_VERSION) and a human-readable build string. */
int32_t     pt_version(void);
const char *pt_build_info(void);

/* Copies the calling thread's last error message into buf (always NUL-terminated).
 * Returns the message length. */
size_t pt_last_error(char *buf, size_t len);

/* Philox4x32-10 stream of the set-up draws (FluctuationSpecularity's per-surface jitter,
 * RT:1151-1157, re-keyed from the runtime's RANDOM_NUMBER): path id 2^62, outside every
 * trace's path range. */
#define PT_SETUP_STREAM 4611686018427387904ull

/* Host-side draws first_draw .. first_draw + n - 1 of stream `path` under `seed`: the same
 * 53-bit uniforms in [0, 1) the kernels draw for path `path` (no GPU needed).  Hosts use it
 * for set-up randomness that must follow the library's streams. */
int pt_philox_fill(uint64_t seed, uint64_t path, uint64_t first_draw, int64_t n, double *out);

/* ---- context: one GPU, one HIP stream, one RNG seed ---------------------------- */
/* Replaces MPI_Init/init_random_seed (main.f90:7-16, RT:3314-3330). */
int pt_ctx_create(int device, uint64_t seed, pt_ctx **out);
int pt_ctx_destroy(pt_ctx *ctx);
int pt_ctx_set_seed(pt_ctx *ctx, uint64_t seed);
/* Safety stop: a path ends (flag PT_PATH_EVENT_LIMIT) after this many events.  The
 * reference loops without bound (RT:1693); default 1e8, at most 715827881 (a path's draw
 * counter is 32 bits; larger limits return PT_E_INVALID).  A stopped path has not ended
 * on either side: every tally counts it as traced and NOT transmitted, and
 * pt_stats.event_limited reports how many there were -- a caller must check it (a
 * nonzero count biases the transmission low). */
int pt_ctx_set_event_limit(pt_ctx *ctx, int64_t limit);
/* on != 0: later trace calls run the instrumented kernel that also fills
 * pt_stats.fp64_ops and philox_blocks (for roofline accounting; slower). */
int pt_ctx_set_op_counting(pt_ctx *ctx, int32_t on);
/* Launch tuning of the persistent trace kernels: paths a wave takes from the work queue
 * at once (chunk; grid, alias and adjoint launches too small to give every wave eight
 * chunks use total / (8 waves) instead), idle lanes that trigger a refill (refill_min,
 * 1..64) and workgroups per CU (0: the occupancy maximum).  Results do not depend on
 * these. */
int pt_ctx_set_tuning(pt_ctx *ctx, int32_t chunk, int32_t refill_min, int32_t blocks_per_cu);
/* Long random walks (accelerated precisions; every trace but the ballistic grid): a path
 * past `events` events, or one of the last few of a wave once the work queue is empty,
 * leaves the persistent kernel for a second kernel that runs it on a whole wave (its ball
 * set-up and scans spread over the lanes, its in-ball events 64 at a time).  At most
 * max_records such paths per launch (the buffer holds total / 256 records, at least 65536,
 * at most max_records; 0: none, every path stays); the rest stay in the persistent kernel.
 * live_blocks workgroups (0..64) of that kernel run beside the persistent one, on a second
 * stream, and take handed-over paths (the slowest first) while the persistent kernel still
 * runs; the persistent kernel gets 2 x live_blocks workgroups fewer.  Not for launches
 * built into HIP graphs (pt_batch_trace_graph) or pt_trace_alias_noalloc, whose tail kernel
 * always runs after the persistent one.
 * Defaults: 500 events, 2^20 records, 16 live workgroups.  Results do not depend on these.
 * Mesh batches (pt_batch_trace), whose events cost more in the persistent kernel, run
 * fastest at about 200 events and 32 live workgroups.
 * No counterpart in the reference, whose OpenMP threads each run one cell's particles to
 * the end (RT:1573). */
int pt_ctx_set_tail(pt_ctx *ctx, int64_t events, int64_t max_records, int32_t live_blocks);

/* ---- geometry ------------------------------------------------------------------ */
/* Uploads vgrain (RT:122) as flattened by the host.  Validates indices; at most
 * 2^30 - 1 surfaces. */
int pt_geom_create(pt_ctx *ctx, const pt_geom_desc *desc, pt_geom **out);
int pt_geom_destroy(pt_geom *geom);

/* Sets surface%specularity of every surface (nsurface values in [0, 1], descriptor
 * order) without rebuilding the geometry: the parameters of the specularity gradient.
 * The reference draws u <= specularity at each reflection (determinSpecular,
 * RT:2225-2235); its values come from param%specularity (RT:1150, 1388). */
int pt_geom_set_specularity(pt_geom *geom, const double *specularity);

/*
 * setupSurfaceGeometry for a polygon (src/RTMCphon.f90:1433-1491): centroid, the
 * averaged (not renormalised) fan normal and the tangent frame, from nvertex >= 3
 * vertices (open loop, xyz triples).  tangent is column-major 3x3.  Host-only helper.
 */
int pt_surface_setup_polygon(int32_t nvertex, const double *vertices,
                             double center[3], double normal[3], double tangent[9]);

/* ---- bulk phonon table ----------------------------------------------------------- */
/* Uploads the table; with vel, mfp and weight present it also builds the alias table. */
int pt_modes_create(pt_ctx *ctx, const pt_modes_desc *desc, pt_modes **out);
int pt_modes_destroy(pt_modes *modes);

/* Vose alias table of n >= 1 non-negative weights (host-only helper; pt_modes_create
 * uses it): mode i is drawn as i0 = floor(u0*n), then i0 if u1 < prob[i0] else
 * alias[i0].  No counterpart in the reference (it enumerates every mode). */
int pt_alias_build(int64_t n, const double *weight, double *prob, int32_t *alias);

/* ---- the trace ---------------------------------------------------------------------- */
/*
 * TransmissionFunction (src/RTMCphon.f90:1544-1604) for one sampling mean free path.
 * For every cell (ik, ib, ia) with ia < nangle, traces np particles exactly as
 * calcTransmit does and writes the transmitted COUNT to
 *   counts[(ib-1) + nb*((ik-1) + nk*(ia-1))]
 * (int64; the reference stores count/np as real(8), RT:1834-1838; counts for
 * ia = nangle stay 0 as in RT:1576).  mfp_nm = 1e10 means ballistic (Matthiessen
 * mode, RT:1686-1690).  vangle: nangle polar angles; NULL computes them as RT:1554-1556.
 * Path p of cell c (c = ((ik-1)*nb + (ib-1))*(nangle-1) + (ia-1), the order of the
 * reference's collapsed loop) uses Philox stream c*np + p.
 * stats may be NULL.
 */
int pt_trace_grid(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes,
                  double mfp_nm, int32_t nangle, const double *vangle, int64_t np,
                  int32_t precision, int64_t *counts, pt_stats *stats);

/*
 * Replay of single paths for parity checks: traces paths first_path .. first_path+n-1
 * of the pt_trace_grid numbering and records each one.  A path is traced even when an
 * earlier particle of its cell failed the start-grain check (pt_trace_grid skips it).
 */
int pt_replay(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes,
              double mfp_nm, int32_t nangle, const double *vangle, int64_t np,
              int64_t first_path, int64_t n, int32_t precision, pt_path_record *records);

/*
 * Trajectory recorder (trace.out, RT:1661-1665 and 1818-1831): traces paths first_path +
 * i * stride (i < n) of the pt_trace_grid numbering, one at a time as pt_replay, and
 * stores the positions the reference writes to trace.out for a traced particle: the
 * start position, then the position at the end of every event-loop iteration.  The
 * first max_points of path i go to points[(i * max_points + k) * 3 + 0..2] (FP64, nm),
 * their total number (also beyond max_points) to npoints[i]; records may be NULL.  The
 * reference traces every interval_trace-th particle of its OpenMP master thread; stride
 * plays that role here on the deterministic path numbering.  Host buffers of
 * n * max_points * 24 B (at most 4 GiB).
 */
int pt_trace_log(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes, double mfp_nm,
                 int32_t nangle, const double *vangle, int64_t np, int64_t first_path,
                 int64_t stride, int64_t n, int32_t precision, int64_t max_points,
                 double *points, int64_t *npoints, pt_path_record *records);

/*
 * Alias-mode trace (an extension of the reference's estimator, whose (k, b, angle)
 * enumeration is pt_trace_grid): npaths paths, path i with Philox stream i; each path
 * draws its mode from the alias table, enters through x = 0 with Lambertian polar
 * angle (cos(theta) = sqrt(u)) and random azimuth, scatters internally with mean free
 * path mfp_m / branch_scale[b] (branch_scale NULL: 1), and is traced by the same event
 * loop as calcTransmit (RT:1693-1824).  Adds per mode m = (ib-1) + nb*(ik-1) the paths
 * drawn and transmitted to mode_paths and mode_trans (int64, caller-allocated nb*nk,
 * accumulated).  kappa/kappa_bulk = (3L/4) * E[T/mfp_m] (host side).
 */
int pt_trace_alias(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes, int64_t npaths,
                   int32_t precision, const double *branch_scale, int64_t *mode_paths,
                   int64_t *mode_trans, pt_stats *stats);

/*
 * pt_trace_alias without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and errors as
 * pt_trace_alias; it calls no hipMalloc/hipFree and allocates no host memory.  Its work
 * space (queue and statistics counters, per-mode tallies, branch scales) is one static
 * device array of 1 MiB per GPU, held from the library's load until the process ends
 * and shared by every call on that GPU: calls that overlap on one GPU (from any thread,
 * context or stream) race and must be serialised by the caller.  At most
 * PT_NOALLOC_MAX_MODES modes (nb * nk) and PT_NOALLOC_MAX_BRANCHES branches; larger
 * tables return PT_E_UNSUPPORTED.  The context must have been created on the GPU
 * (pt_ctx_create looks up the array's device address).
 */
#define PT_NOALLOC_MAX_MODES 65536
#define PT_NOALLOC_MAX_BRANCHES 64
int pt_trace_alias_noalloc(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes,
                           int64_t npaths, int32_t precision, const double *branch_scale,
                           int64_t *mode_paths, int64_t *mode_trans, pt_stats *stats);

/* Per-path records of alias-mode paths first .. first+n-1 (parity checks). */
int pt_replay_alias(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes,
                    const double *branch_scale, int64_t first, int64_t n, int32_t precision,
                    pt_path_record *records);

/*
 * Adjoint (an extension; the reference has no derivatives): the alias-mode trace of
 * pt_trace_alias plus likelihood-ratio scores, in one pass with nothing stored per path.
 * With x = T / mfp_m for a path of mode m (T = 1 if transmitted), kappa = c * E[x] with
 * c = kappa_bulk * 3L/4, and for each parameter theta
 *     d kappa / d theta = c * E[(x - baseline) * d log P(path) / d theta],
 * P(path) the path's probability: every specular decision on a surface of group g
 * contributes 1/p (specular) or -1/(1-p) (diffuse), and the free flights of branch b
 * (rate branch_scale[b] / mfp_m) contribute n_scat / branch_scale[b] - L / mfp_m
 * (n_scat internal scatterings, L the path length).  Any constant baseline keeps the
 * estimate unbiased; a value near E[x] lowers its variance.
 * surface_group: nsurface entries, the gradient group of each surface or -1 (several
 *   surfaces may share a group: the derivative is then with respect to a common shift
 *   of their specularities); grouped surfaces need 0 < p < 1.  ngroup <= 32.
 * batch_sums: nbatch x (3 + ngroup + nbranch) doubles, accumulated; path i adds to batch
 *   i mod nbatch the values [1, x, x^2, (x - baseline) * score_0, ...], the scores
 *   ordered as the ngroup specularity groups, then the nbranch branch scales.
 * mode_paths / mode_trans are accumulated as by pt_trace_alias.
 */
int pt_adjoint_alias(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes, int64_t npaths,
                     int32_t precision, const double *branch_scale,
                     const int32_t *surface_group, int32_t ngroup, double baseline,
                     int32_t nbatch, int64_t *mode_paths, int64_t *mode_trans,
                     double *batch_sums, pt_stats *stats);

/*
 * Alias-mode trace with per-grain heat-flux tallies (an extension; the reference tallies
 * only the transmitted count per cell, RT:1833-1838).  As pt_trace_alias, and for every
 * grain g = 0 .. ngrain_total-1 adds over all flights spent in grain g (whichV during the
 * flight) of every path of mode m:
 *     grain_tally[2*g]   += dx / mfp_m     (signed x displacement: the heat flux)
 *     grain_tally[2*g+1] += l  / mfp_m     (path length)
 * in FP64.  A path's displacements sum to L T (up to the 1e-7 nm nudges, RT:1779), so
 * sum_g grain_tally[2g] = L sum_paths T / mfp_m: kappa/kappa_bulk = (3/4) sum_g
 * grain_tally[2g] / npaths splits into per-grain shares.
 */
int pt_trace_alias_grains(pt_ctx *ctx, const pt_geom *geom, const pt_modes *modes,
                          int64_t npaths, int32_t precision, const double *branch_scale,
                          int64_t *mode_paths, int64_t *mode_trans, double *grain_tally,
                          pt_stats *stats);

/* ---- batches of mesh structures (an extension) ------------------------------------ */
/*
 * Many PT_GRAIN_MESH structures traced in one persistent launch: the structures share
 * the box (size, scopes, wall laws and materials) and are merged into one geometry with
 * disjoint grain sets; structure s owns paths [sum_{r<s} npaths[r], ... + npaths[s]) of
 * the alias-mode numbering (Philox stream = path id), so each structure's results equal
 * those of pt_trace_alias on its own geometry with those path ids.  Each grain gets a
 * BVH over its candidate surfaces (leaf_size prims per leaf at most) that replaces the
 * linear scan of findCollisionSurface (RT:1989-2096) with the same result; each structure
 * gets a grid x grid locator of its incident facets for the start grain.
 */
typedef struct pt_batch pt_batch;
int pt_batch_create(pt_ctx *ctx, int32_t nstruct, const pt_geom_desc *const *descs,
                    const int64_t *npaths, int32_t leaf_size, int32_t grid, pt_batch **out);
int pt_batch_destroy(pt_batch *batch);
/* info: [surfaces, grains, BVH nodes, BVH prims, BVH depth, paths] of the merged batch */
int pt_batch_info(const pt_batch *batch, int64_t *info);
/* every facet of structure s gets specularity mean[s] (0 < mean[s] < 1); box walls keep
 * theirs.  The gradient of pt_batch_trace is with respect to these values. */
int pt_batch_set_specularity(pt_batch *batch, const double *mean);
/* Top BVH levels staged in each workgroup's LDS (0: none, at most 3; a tuning knob, no
 * reference counterpart): a workgroup copies the top `levels` levels of the grain BVHs of
 * the structure its waves are tracing into LDS, and its scans of that structure read those
 * nodes there.  Results do not depend on it; the LDS it takes (2^levels - 1 nodes of 144
 * bytes per grain of the largest structure) can lower the workgroups per CU. */
int pt_batch_set_top_levels(pt_batch *batch, int32_t levels);
/*
 * One launch over the whole batch with the context's seed.  Per structure s, adds to
 * sums[5 s .. 5 s + 4]: paths, sum x, sum x^2, sum (x - baseline[s]) * score_s and the sum
 * of its squares (for the gradient's standard error), with
 * x = T / mfp_m (kappa/kappa_bulk = (3L/4) E[x]) and score_s the likelihood-ratio score
 * of the structure's mean specularity (1/p per specular, -1/(1-p) per diffuse decision on
 * its facets; baseline may be NULL for 0).  finish_ms[s]: time from the kernel's start to
 * the end of the structure's last path (ms), or NULL.  precision: PT_PREC_*.
 */
int pt_batch_trace(pt_ctx *ctx, pt_batch *batch, const pt_modes *modes, int32_t precision,
                   const double *baseline, double *sums, double *finish_ms, pt_stats *stats);
/* As pt_batch_trace, through a hipGraph of (zero fill, trace) built on the first call for
 * these modes and precision and replayed afterwards with the context's current seed. */
int pt_batch_trace_graph(pt_ctx *ctx, pt_batch *batch, const pt_modes *modes,
                         int32_t precision, const double *baseline, double *sums,
                         double *finish_ms, pt_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* PTRANS_H */
