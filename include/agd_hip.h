/* agd_hip.h -- C ABI of the MI355X (gfx950) eikonal solver library
 * (libagd_hip.so).
 *
 * The library replaces the device side of the reference's GPU eikonal path:
 * the NVRTC-compiled `Update` kernel and the cupy host loops that drive it.
 * Every entry point takes plain pointers and sizes; device pointers are
 * caller-owned buffers in HBM (the library only owns its internal
 * workspaces).  All work is enqueued on the stream passed in (a hipStream_t,
 * passed as void*; NULL = default stream).  Functions return 0 on success and
 * a negative agd_status otherwise; agd_last_error() describes the failure.
 *
 * Data layout ("tiled"): a grid of shape shape[0..ndim) is padded up to a
 * multiple of the tile shape shape_i and stored tile after tile:
 *   n_t = n_o * size_i + n_i,
 * n_o the row-major index of the tile in the tile grid shape_o, n_i the
 * row-major index of the cell inside its tile (last axis fastest).  Padded
 * cells carry +inf (values, rhs, geometry).  Multi-component fields
 * (geometry, flow) are component-major: field[k * size_tot + n_t].
 * This is the reference's bi-level layout (Update.h:60-66, Grid.h:15-35,
 * FiniteDifferences.py:164-222).
 */
#ifndef AGD_HIP_H
#define AGD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ enums */
enum agd_status {
  AGD_OK = 0,
  AGD_EINVAL = -1,  /* invalid argument (shape, model, tile, pointer)      */
  AGD_EHIP = -2,    /* a HIP runtime call failed                            */
  AGD_NOCONV = -3,  /* nitermax_o reached before convergence                */
  AGD_ENOMEM = -4,  /* workspace allocation failed                          */
};

enum agd_model {
  AGD_ISOTROPIC = 0, /* Isotropic_.h: unit offsets, weights h^-2 (iso_weights) */
  AGD_DIAGONAL = 1,  /* Isotropic_.h with geometry: per-cell weights        */
  AGD_RIEMANN = 2,   /* Riemann_.h: Selling decomposition of the dual metric */
  AGD_RANDER = 3,    /* Riemann_.h + drift (FiniteDifferences.h:41-45)      */
  AGD_TTI = 4,       /* TTI_.h: min/max of 7 Selling schemes; geometry
                        [l0, l1, q00, q10, q11, T (d x d)]; precomputed scheme
                        only (agd_eikonal_scheme)                            */
  /* Curvature-penalised models on the (x, y, theta) grid, ndim 3, theta
   * periodic; geometry [1/xi, kappa, cos theta, sin theta] per cell in grid
   * units (_SetGeometry.py:83-109); precomputed scheme only.  Each sub-scheme
   * stores its offsets as pairs of codes (x - e, x + e); a forward offset
   * (one-sided, x + e only) has its x - e code set to -1. */
  AGD_REEDSSHEPP2 = 5,       /* ReedsShepp2.h: Selling of the relaxed v v^T, 6 symmetric offsets */
  AGD_REEDSSHEPPFWD2 = 6,    /* ReedsSheppForward2.h: 1 symmetric (0,0,1) + 6 forward offsets */
  AGD_ELASTICA2 = 7,         /* Elastica2.h: 5 Fejer directions x 6 forward offsets */
  AGD_DUBINS2 = 8,           /* Dubins2.h: min of 2 schemes of 6 forward offsets */
};

enum agd_scalar { AGD_F32 = 0, AGD_F64 = 1 };

enum agd_solver_kind {
  AGD_AGSI = 0,   /* adaptive Gauss-Seidel: tag self + faces on change
                     (_solvers.py:122-193, Propagation.h:160-164)          */
  AGD_FIM = 1,    /* fast iterative method with front score
                     (_solvers.py:94-120, Propagation.h:127-159)           */
  AGD_GLOBAL = 2, /* every tile on every pass (_solvers.py:77-92)          */
};

/* ------------------------------------------------------------ descriptors */
typedef struct agd_eikonal_desc {
  int32_t model;          /* agd_model                                   */
  int32_t ndim;           /* 2 or 3                                      */
  int32_t scalar;         /* agd_scalar                                  */
  int32_t solver;         /* agd_solver_kind                             */
  int64_t shape[3];       /* grid shape (unused axes = 1)                */
  int32_t shape_i[3];     /* tile shape; supported: see agd_tile_supported */
  int32_t periodic[3];    /* per-axis periodic boundary                  */
  int32_t niter_i;        /* in-tile sweeps per tile visit               */
  int32_t fim_front_width;/* FIM score of a freshly stabilised tile (>=2) */
  double atol, rtol;      /* |u_old-u_new| <= atol + rtol |u_new|        */
  int64_t nitermax_o;     /* maximum number of grid-level passes         */
  double iso_weights[3];  /* AGD_ISOTROPIC: per-axis weights (h^-2)      */
  int32_t check_interval; /* passes between host convergence checks (>=1);
                             with use_graph, a multiple of 6 lets every pass
                             read only its own row of worklist counters */
  int32_t activate_all;   /* 1: start with every tile active (warm start) */
  int32_t time_launches;  /* 1: time every pass with an event pair (kernel_ms) */
  int32_t use_graph;      /* 1: capture check_interval passes once as a hipGraph and replay it */
  int32_t order;          /* 1 (or 0) or 2: second-order upwind differences
                             where the solution is smooth (HFM.h:91-134) */
  int32_t convex;         /* curvature models ReedsSheppForward2, Elastica2,
                             Dubins2: 1 for their Convex variants, whose
                             vehicle turns left only (kernel_traits.py:129-130) */
  double order2_threshold;/* order 2: |u-2v+v2| < threshold * max(u-v) (0.3) */
  /* Source factorization (EuclideanFactor.h:8-50, Constants.h:50-61,
   * TTI_.h:139-212): the cells with |x - factor_origin|^2 < factor_radius2
   * (grid units, Euclidean) shift their neighbour values by the
   * finite-difference error of T(x) = |x - factor_origin|_M, M = factor_metric.
   * 0 disables it.  Not with batched solves, nor with the precomputed scheme
   * of Riemann (TTI's precomputed scheme takes it: a table of the ball's
   * shifts is computed at the start of every solve, in a workspace the
   * allocation-free entry point does not have). */
  double factor_radius2;
  double factor_origin[3];/* seed position in grid units                 */
  double factor_metric[14];/* M in grid units: squared costs (Isotropic, Diagonal:
                             ndim entries), lower triangle (d00, d10, d11,
                             d20, d21, d22) for Riemann and Rander, or for TTI
                             [l0, l1, q00, q10, q11, A (ndim x ndim, row-major)]
                             with T(x) = N0(A x) (_Kernel.py:301-308) */
  /* bound_active_blocks (_solvers.py:145-181,195-224, Propagation.h:53-99):
   * with the AGSI solver, a tile whose smallest changed value is at or above
   * a moving threshold is frozen (kept in the list without propagating)
   * until the threshold passes it; the threshold is advanced after every
   * pass to keep about bound_active_blocks tiles active.  0 disables it. */
  int64_t bound_active_blocks;
  double minchg_delta_min; /* smallest threshold increase (> 0)           */
  /* Curvature models: relaxation eps^2 of the self outer products of the
   * directions of motion (Decomp_v_.h, _Kernel.py:318-321; eps = 0.1 gives
   * 0.01), > 0; unused by the other models. */
  double curvature_relax;
} agd_eikonal_desc;

/* Device pointers, tiled layout.  Scalars are float (AGD_F32) or double. */
typedef struct agd_eikonal_io {
  void* values;               /* in/out: initial supersolution, then u    */
  const void* geom;           /* ngeom components (NULL for isotropic)    */
  const void* rhs;            /* cost; seed value on seeds; +inf in walls */
  const uint32_t* seeds;      /* 1 bit per cell (bit n_t%32 of word n_t/32): fixed cells */
  const uint32_t* trigger;    /* 1 bit per cell: cells whose tiles start active (NULL: seeds) */
  const uint8_t* wall_dist;   /* saturated L1 distance to walls, 255 = far; NULL = no walls */
  /* Precomputed scheme (both NULL: the stencil is rebuilt at every tile
   * visit).  Filled by agd_eikonal_scheme from geom and wall_dist.  Each
   * component is one tiled field of size_tot entries:
   *   Riemann: scheme_weights nsym components of the scalar type,
   *            scheme_codes 2*nsym int32 components (neighbour x - e_k then
   *            x + e_k, resolved against the tiled grid: >= 0 tiled index,
   *            <= -2 slot -2-code of the same tile, -1 none);
   *   TTI:     weights 7*nsym components (sub-scheme m at m*nsym), codes
   *            7*2*nsym components (sub-scheme m at m*2*nsym), then one
   *            component holding the min/max flag, then, with order 2, 7*2*nsym
   *            components: the codes of the further neighbours x -+ 2 e_k.
   * agd_scheme_lengths gives the component counts. */
  void* scheme_weights;
  int32_t* scheme_codes;
  /* Entries (not bytes) available in scheme_weights and scheme_codes; the
   * solve and agd_eikonal_scheme refuse buffers shorter than the layout
   * above needs for the descriptor's model, dimension and order. */
  int64_t scheme_weights_len;
  int64_t scheme_codes_len;
} agd_eikonal_io;

typedef struct agd_eikonal_stats {
  int64_t niter_o;        /* grid-level passes with a non-empty tile list */
  int64_t tile_updates;   /* tile visits summed over passes               */
  int64_t cell_updates;   /* tile_updates * size_i                         */
  int64_t launches;       /* kernel launches issued (incl. empty passes)  */
  double gpu_ms;          /* HIP-event time of the solve loop             */
  int32_t converged;      /* 1 if the tile list emptied                   */
  int32_t grid_blocks;    /* workgroups per launch                         */
  double kernel_ms;       /* sum of per-pass kernel times (time_launches)  */
} agd_eikonal_stats;

/* Flow export (reference: Update.h:171-209 compiled with FLOW_* macros,
 * driven by _PostProcess.py:46-95).  Any output may be NULL. */
typedef struct agd_flow_io {
  void* flow_weights;     /* [nsym][size_tot]  w_k (u - v_k)_+           */
  void* flow_weightsum;   /* [size_tot]                                  */
  int32_t* flow_indices;  /* [nsym][size_tot] tiled index of the upwind neighbour, INT32_MAX if weight 0 */
  void* flow_vector;      /* [ndim][size_tot] sum_k fw_k (+-e_k)          */
} agd_flow_io;

/* ------------------------------------------------------------ library */
/* The message of the last failure on the calling thread. */
const char* agd_last_error(void);
int agd_version(void);
/* 1 if (model, ndim, shape_i) has a compiled kernel. */
int agd_tile_supported(int32_t model, int32_t ndim, const int32_t* shape_i);
/* 1 if (model, ndim, shape_i) has a compiled batched kernel: the shapes above
 * and, for agd_eikonal_solve_batched only, the 2D tiles of fewer than 64 cells
 * (4x4, 4x8, 8x4; the reference takes any shape_i, kernel_traits.py:55-90). */
int agd_batch_tile_supported(int32_t model, int32_t ndim, const int32_t* shape_i);
int agd_device_count(int32_t* count);

/* ------------------------------------------------------------ solver handles
 * A handle owns the workspace of the solves run through it: the device
 * worklist (tile stamps, sub-lists, counters, pass histogram, FIM scores),
 * pinned host counters, events, a private capture stream and the last
 * captured chunk of passes, which a later solve with the same kernel and
 * arguments replays instead of capturing again.  It is bound to the device
 * current at creation and to `stream`.  Handles share no state: solves
 * through different handles may run at the same time (on different streams
 * or host threads).  One handle serves one host thread at a time.
 * (The reference keeps its solver state in module __constant__ globals,
 * cupy_module_helper.py:35-49, and reruns its host loop per call,
 * _solvers.py:122-193.)  desc may be NULL; given, the workspace is sized
 * for its grid.  Workspaces grow on demand. */
typedef struct agd_solver agd_solver;
int agd_solver_create(const agd_eikonal_desc* desc, void* stream, agd_solver** out);
int agd_solver_destroy(agd_solver* solver);
/* agd_eikonal_solve, agd_eikonal_solve_batched and agd_eikonal_check through
 * a handle, on its stream. */
int agd_solver_solve(agd_solver* solver, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     agd_eikonal_stats* stats);
typedef struct agd_batch_instance agd_batch_instance;
int agd_solver_solve_batched(agd_solver* solver, const agd_eikonal_desc* desc, int32_t ninst,
                             const agd_batch_instance* instances, const agd_eikonal_io* io,
                             agd_eikonal_stats* stats);
int agd_solver_check(agd_solver* solver, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     double atol, double rtol, int64_t* violations);
/* Cells a solve left above their scheme update (no counterpart in the
 * reference, whose schedule has the same gap): counts the non-seed cells whose
 * update lies below their value by more than atol + rtol |u|, and sets their
 * bits in stale_bits (tiled, 1 bit per cell; accumulated into, not cleared).
 * A solve with io->trigger = stale_bits continues from them.  A changed tile
 * re-lists only its face neighbours (Propagation.h:18-35), so a stencil that
 * reaches two tiles away (the curvature models') can leave such cells behind. */
int agd_solver_check_stale(agd_solver* solver, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                           double atol, double rtol, uint32_t* stale_bits, int64_t* stale);
/* Turns io->values into a supersolution of the scheme before a warm-started
 * solve (no counterpart in the reference): each round counts the non-seed
 * cells whose scheme update exceeds their value by more than atol + rtol |u|
 * (agd_eikonal_check's test) and raises them to +inf.  A start with no such
 * cell lies above the discrete fixed point (comparison principle of the
 * monotone scheme), so the decreasing iteration converges from it to the
 * cold-start solution.  Stops after a round that counts nothing, or after
 * max_rounds rounds; *violations_left = the last round's count (0: the
 * start is a supersolution), *rounds = rounds run (one grid pass each). */
int agd_solver_repair(agd_solver* solver, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                      double atol, double rtol, int32_t max_rounds, int64_t* violations_left,
                      int32_t* rounds);

/* Solve the eikonal equation in place on `values` (replaces Solve('eikonal'),
 * _solvers.py:15-57, and the Update kernel, Update.h:20-227).  Returns
 * AGD_NOCONV (and still fills stats) if nitermax_o passes did not converge.
 * Without a handle, the call allocates its own workspace and frees it
 * before returning (re-entrant; no state survives the call). */
int agd_eikonal_solve(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                      agd_eikonal_stats* stats, void* stream);

/* agd_eikonal_solve without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under stream capture, tight serving loops or harnesses that forbid
 * hipMalloc.  The worklist lives in a fixed 12 MiB device array of the
 * library (one per device, held until the process ends), which bounds the
 * grid to about 150,000 tiles (200^3 cells in 4x4x4 tiles, 3000^2 in 8x8)
 * and nitermax_o to about 2,000 passes; larger solves are refused with
 * AGD_EINVAL.  Passes launch eagerly and stats->gpu_ms is 0 (no events).
 * Every call on a device shares that device's array, so calls that overlap
 * on one device (several streams or threads) race: serialise them. */
int agd_eikonal_solve_noalloc(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, void* stream);

/* One global pass exporting the upwind flow (replaces Solve('flow'),
 * _PostProcess.py:79, Update.h:171-209).  Updates `values` like the
 * reference's flow pass (DECREASING rule).  */
int agd_eikonal_flow(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     const agd_flow_io* flow, void* stream);

/* One instance of a batched solve (no counterpart in the reference, whose
 * solver runs one grid per call).  Instances share the descriptor's model,
 * tile shape, scalar type and solver; each has its own shape, isotropic
 * weights and tolerances. */
struct agd_batch_instance {
  int64_t shape[3];        /* instance grid (unused axes = 1)            */
  double iso_weights[3];   /* AGD_ISOTROPIC: per-axis weights (h^-2)     */
  double atol, rtol;       /* convergence tolerances of the instance     */
};

/* Solve several instances in one worklist (one pass launch per grid pass for
 * the whole batch).  io buffers are the concatenation, in instance order, of
 * each instance's tiled arrays (values, rhs, geom component-major over the
 * concatenated cells, seed and trigger bits).  desc->shape is ignored; no
 * walls, no periodic axes, no precomputed scheme; FP32 only.  Converges when
 * every instance has converged. */
int agd_eikonal_solve_batched(const agd_eikonal_desc* desc, int32_t ninst,
                              const agd_batch_instance* instances, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, void* stream);

/* Residual check ("the solver as judge" of a warm start; no counterpart in
 * the reference): counts the non-seed cells whose scheme update exceeds their
 * value by more than atol + rtol |u|.  After a solve started from a
 * supersolution the count is 0; a positive count means the start was below
 * the fixed point somewhere (the decreasing iteration cannot recover). */
int agd_eikonal_check(const agd_eikonal_desc* desc, const agd_eikonal_io* io, double atol,
                      double rtol, int64_t* violations, void* stream);

/* Precompute the scheme of every cell into io->scheme_weights /
 * io->scheme_codes (replaces the precompute_scheme export kernel,
 * _Kernel.py:254-268,346-359, Update.h:76-80 import_scheme_macro).  Needs
 * geom (and wall_dist if walls are present).  Riemann and TTI, as
 * agd_scheme_supported reports.  TTI stores its 7 sub-schemes: weights
 * 7*nsym components, codes 7*2*nsym components (sub-scheme m at m*2*nsym)
 * plus one component, 1 where the update is the min of the sub-schemes'
 * updates and 0 where it is their max (TTI_.h:127); with desc->order 2,
 * 7*2*nsym more components follow: the codes of the further neighbours
 * x -+ 2 e_k of the second-order scheme (HFM.h:91-134), which
 * agd_eikonal_solve then reads. */
int agd_scheme_supported(int32_t model, int32_t ndim, const int32_t* shape_i);
/* Components of scheme_weights and scheme_codes for (model, ndim, order). */
int agd_scheme_lengths(int32_t model, int32_t ndim, int32_t order, int64_t* weight_components,
                       int64_t* code_components);
int agd_eikonal_scheme(const agd_eikonal_desc* desc, const agd_eikonal_io* io, void* stream);

/* Reverse-mode sensitivities (replaces SolveAD's reverse branch,
 * _PostProcess.py:194-219: graph_reverse.py:10-77 + Kernel_LinearUpdate.h).
 * Solves (D - W)^T lambda = sens for nrhs right-hand sides with the tile
 * Gauss-Seidel solver, where W is the flow-weight graph (flow_weights,
 * flow_indices as produced by agd_eikonal_flow) and D = flow_weightsum
 * (1 where it is 0).  sens and lambda: [nrhs][size_tot] tiled; lambda is
 * overwritten.  atol/rtol as in _PostProcess.py:126-138. */
int agd_eikonal_adjoint(const agd_eikonal_desc* desc, const void* values,
                        const void* flow_weights, const int32_t* flow_indices,
                        const void* flow_weightsum, int32_t nrhs, const void* sens,
                        void* lambda, double atol, double rtol,
                        agd_eikonal_stats* stats, void* stream);

/* Forward-mode variation of the solution (replaces SolveAD's forward
 * branch, _PostProcess.py:181-191): solves (D - W) dvalues = rhs with the
 * same flow graph, untransposed; rhs = cost * dcost off the seeds and the
 * seed value variation on the seeds.  rhs and dvalues: [nrhs][size_tot]
 * tiled; dvalues is overwritten. */
int agd_eikonal_forward_ad(const agd_eikonal_desc* desc, const void* values,
                           const void* flow_weights, const int32_t* flow_indices,
                           const void* flow_weightsum, int32_t nrhs, const void* rhs,
                           void* dvalues, double atol, double rtol,
                           agd_eikonal_stats* stats, void* stream);

/* ------------------------------------------------------------ geodesics */
/* Geodesic backtracking from tips (Kernel_GeodesicODE.h:265-413 with the
 * imported flow of GeodesicODE_Opt.h:222-238; driven by
 * _GetGeodesics.py:86-131, which launches the reference's GeodesicODE
 * kernel).  One geodesic per lane, midpoint steps along the normalised,
 * causally filtered interpolation of the upwind flow vector. */
typedef struct agd_geodesic_desc {
  int32_t ngeo;            /* number of geodesics                          */
  int32_t max_len;         /* points per geodesic row of x (>= 1)          */
  int32_t eucl_delay;      /* PastSeed history delay, in steps             */
  int32_t nymin_delay;     /* Stationnary history delay, in steps          */
  int32_t pastseed_delay;  /* steps before the PastSeed test is enabled    */
  int32_t chart_ndim_s;    /* charts: axes of chart_mapping (0: no charts) */
  double step;             /* ODE step, in grid units (pixels)             */
  double causality_tolerance; /* corner exclusion: value >= min + tol / flow weight sum */
  /* Charts (GeodesicODE_Opt.h:43-99, Kernel_GeodesicODE.h:317-320, 406):
   * after a step taken from a cell whose eucl entry is AGD_EUCL_CHART, the
   * last chart_ndim_s coordinates of the point are replaced by the
   * multilinear interpolation of chart_mapping (layout of agd_chart_paste;
   * corners clamped to the grid), or by the mapping of the nearest corner
   * when the mapped corners' summed variance reaches chart_jump_variance
   * (+inf: never).  A mapping with a NaN leaves the point unchanged. */
  const void* chart_mapping;
  double chart_jump_variance;
} agd_geodesic_desc;

/* eucl value marking the cells where geodesics jump to another chart; the
 * capped distances stay below it (255 = beyond the bound). */
#define AGD_EUCL_CHART 254

/* Capped distance to the seeds read by the PastSeed criterion
 * (_GetGeodesics.py:50-67, inf_convolution.py:24-93): eucl = 0 on the seed
 * cells (bit array `seeds`, tiled), then `niter` min-plus passes over the 3^d
 * neighbourhood with the integer kernel round(mult |offset|) saturating at
 * 255, and values above `bound` set to 255.  eucl and tmp: tiled uint8
 * arrays of the solver's padded size (padding = 255). */
int agd_geodesic_eucl(const agd_eikonal_desc* desc, const uint32_t* seeds, int32_t niter,
                      int32_t mult, int32_t bound, uint8_t* eucl, uint8_t* tmp, void* stream);
/* Backtracks gd->ngeo geodesics.  values, flow_vector (ndim components,
 * grid units, towards the upwind neighbours), flow_weightsum: tiled arrays of
 * desc->scalar from agd_eikonal_flow; eucl from agd_geodesic_eucl.
 * x: ngeo x max_len x ndim scalars, row start = starting point (continuous
 * grid index); points 1 .. len-1 of each row are written.  stop: 0 continue
 * (max_len reached), 1 AtSeed, 2 InWall, 3 Stationnary, 4 PastSeed,
 * 5 VanishingFlow.  1 + max(eucl_delay, nymin_delay) must be <= 256. */
int agd_geodesic_ode(const agd_eikonal_desc* desc, const agd_geodesic_desc* gd, const void* values,
                     const void* flow_vector, const void* flow_weightsum, const uint8_t* eucl,
                     void* x, int32_t* len, int8_t* stop, void* stream);
/* The same, with the flow at each cell the ODE enters recomputed from the
 * scheme instead of read from exported arrays (geodesic_online_flow,
 * GeodesicODE_Opt.h:101-238, _GetGeodesics.py:92-93): desc and io as for
 * agd_eikonal_flow (values, geom, rhs, seeds, walls; the model, order and
 * source factorization of desc), no flow buffers.  Tiles of 64 cells or more. */
int agd_geodesic_ode_online(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                            const agd_geodesic_desc* gd, const uint8_t* eucl, void* x, int32_t* len,
                            int8_t* stop, void* stream);

/* ------------------------------------------------------------ charts */
/* Value pasting for manifolds described by local charts (ChartPaste.h:71-167,
 * run by the stopping criterion _StopCriterion.py:60-80 after each converged
 * solve, up to chart_nitermax times; _StopCriterion.py:93-165 prepares the
 * mapping).  For every cell x, mapping gives, for the last ndim_s axes, the
 * continuous grid index (cell centre = integer) of the point identified with
 * x; the first ndim - ndim_s coordinates are kept.  The values there are the
 * multilinear interpolation over the 2^ndim_s surrounding cells; points within
 * 1e-4 of a cell of the last layer are moved onto it, points beyond and NaN
 * mappings are skipped; the axes are not wrapped (as the reference's paste).
 * Where the interpolated value is smaller by more than atol + rtol |value|
 * (0, 0: by any amount, as the reference) the cell takes it, its trigger bit
 * is set and *count (a device int32) is incremented.  Evaluation and update are
 * two passes, so the result does not depend on the order of the cells.
 * values: tiled, desc->scalar.  mapping: cell-major [ndim_s][n_s] in
 * desc->scalar, n_s = product of the last ndim_s dimensions (broadcast over
 * the first ones).  scratch: size_tot elements of desc->scalar.  trigger
 * (tiled bits) and *count are accumulated into, not cleared.  A solve with
 * io->trigger = trigger then continues from the pasted values. */
int agd_chart_paste(const agd_eikonal_desc* desc, void* values, const void* mapping, int32_t ndim_s,
                    double atol, double rtol, void* scratch, uint32_t* trigger, int32_t* count,
                    void* stream);

/* ------------------------------------------------------------ layout */
/* block_expand (FiniteDifferences.py:164-195): src cell-major
 * [ncomp][shape], dst tiled [ncomp][size_tot], padded with pad.
 * elem_size: 1, 4 or 8 bytes.  src_cell_stride = 0 broadcasts one value per
 * component to the whole grid. */
int agd_block_expand(int32_t ndim, const int64_t* shape, const int32_t* shape_i,
                     int32_t ncomp, int32_t elem_size, const void* src,
                     int64_t src_cell_stride, void* dst, double pad, void* stream);
/* block_squeeze (FiniteDifferences.py:197-222): inverse of the above. */
int agd_block_squeeze(int32_t ndim, const int64_t* shape, const int32_t* shape_i,
                      int32_t ncomp, int32_t elem_size, const void* src, void* dst,
                      void* stream);

/* Geometry preparation (_SetGeometry.py:127-142 for Riemann / Rander,
 * riemann.py:38-40,68-73, rander.py:61-68, misc.py:8-15): from a cell-major
 * metric m[d][d][cells] (and drift w[d][cells] for Rander), grid scales h[d],
 * write the tiled geometry of the solver:
 *   Riemann: flat((diag(h) m diag(h))^-1)                    (symdim comps)
 *   Rander : [flat((diag(h) m diag(h))^-1), h * w]           (symdim + d)
 *   Diagonal: m holds costs[d][cells]; geometry = (costs*h)^-2 (d comps)
 *   TTI    : m holds [l0, l1, q00, q10, q11][cells], w the transformation
 *            A[d][d][cells] (NULL: identity); geometry
 *            [l, flat(q), (A diag(h))^-T row-major]             (5 + d*d)
 *            (Seismic/tti.py:113-125)
 * The inverse is computed in double precision.  cell_stride = 0 broadcasts
 * a constant metric.  dst element type follows `scalar`. */
int agd_prepare_geometry(int32_t model, int32_t ndim, const int64_t* shape,
                         const int32_t* shape_i, int32_t scalar, const double* h,
                         const float* m, const float* w, int64_t cell_stride,
                         void* geom, void* stream);

/* Seeds (_SetArgs.py:116-167): for each seed index (tiled cell index),
 * set rhs = seed value, set the seed bit and the trigger bit.  Bits are
 * OR-ed in, so callers zero the bit arrays first. */
int agd_scatter_seeds(int32_t scalar, int64_t nseeds, const int64_t* tiled_index,
                      const double* seed_values, void* rhs, uint32_t* seeds,
                      uint32_t* trigger, void* stream);

/* Walls (_SetGeometry.py:262-282, _SetArgs.py:124-126): wall cells get
 * rhs = +inf and their seed bit; wall_dist = saturated L1 distance to the
 * walls (computed by `bound` min-plus passes, Kernel_InfConvolution.h),
 * 255 beyond `bound`.  walls: tiled uint8 mask (padding = 0). */
int agd_prepare_walls(const agd_eikonal_desc* desc, const uint8_t* walls, int32_t bound,
                      void* rhs, uint32_t* seeds, uint8_t* wall_dist, void* stream);

/* Debug export of the per-cell stencil (weights, offsets) computed by the
 * device code from a component-major geometry of n cells (cell-major, not
 * tiled): weights[n][nsym], offsets[n][nsym][ndim].  Used to check that the
 * Selling offsets are bit-identical to the oracle's
 * (Geometry2.h:40-48, Geometry3.h:43-58). */
int agd_stencil_export(int32_t model, int32_t ndim, int32_t scalar, int64_t n,
                       const void* geom, const double* iso_weights, void* weights,
                       int32_t* offsets, void* stream);

/* ------------------------------------------------------------ memory */
int agd_malloc(void** ptr, int64_t bytes);
int agd_free(void* ptr);
int agd_memcpy_htod(void* dst, const void* src, int64_t bytes, void* stream);
int agd_memcpy_dtoh(void* dst, const void* src, int64_t bytes, void* stream);
int agd_memset(void* dst, int32_t value, int64_t bytes, void* stream);
int agd_stream_synchronize(void* stream);

/* ------------------------------------------------------------ surrogate
 * Kernels of the convolutional surrogate that warm-starts the iteration (no
 * counterpart in the reference).  Activations are BF16 NHWC with channels
 * padded to multiples of 32 (W a multiple of 16 for the forward
 * convolution); weights BF16 [9][cout][cin], tap = 3*(dy+1)+(dx+1);
 * accumulation in FP32 on the matrix cores. */
int agd_nn_conv3x3(const void* in, const void* w, void* out, int32_t N, int32_t H, int32_t W,
                   int32_t Cin, int32_t Cout, int32_t relu, void* stream);
/* dw[9][cout][cin] (FP32) += gradient of the convolution w.r.t. its weights. */
int agd_nn_conv3x3_wgrad(const void* in, const void* gout, float* dw, int32_t N, int32_t H,
                         int32_t W, int32_t Cin, int32_t Cout, void* stream);
/* Weights of the data gradient: gin = agd_nn_conv3x3(gout, wt) with wt from
 * this call (taps reversed, cin and cout exchanged). */
int agd_nn_conv3x3_flip(const void* w, void* wt, int32_t Cin, int32_t Cout, void* stream);
int agd_nn_relu_mask(void* g, const void* x, int64_t n, void* stream);
/* 2x2 average pooling / nearest 2x upsampling (H, W: the coarse grid) and
 * their adjoints (backward = 1). */
int agd_nn_pool2(const void* in, void* out, int32_t N, int32_t H, int32_t W, int32_t C,
                 int32_t backward, void* stream);
int agd_nn_up2(const void* in, void* out, int32_t N, int32_t H, int32_t W, int32_t C,
               int32_t backward, void* stream);
int agd_nn_add(const void* a, const void* b, void* out, int64_t n, void* stream);
/* Masked mean square error on channel 0 of pred, scaled by 1/scale^2 and
 * divided by count; writes the gradient and adds the loss to *loss. */
int agd_nn_mse(const void* pred, const float* target, const float* mask, void* gpred,
               int64_t npix, int32_t C, double scale, double count, float* loss, void* stream);
int agd_nn_adam(float* w, const float* g, float* m, float* v, void* wb, int64_t n, double lr,
                double b1, double b2, double eps, int32_t step, void* stream);
/* Warm-start values on the fine H x W grid: inflate * scale * (bilinear
 * upsampling by `factor` of channel 0 of the coarse prediction, clamped below
 * at 0); scale (H x W, FP32) may be NULL. */
int agd_nn_warm_values(const void* pred, int32_t Hc, int32_t Wc, int32_t C, int32_t factor,
                       double inflate, const float* scale, float* out, int32_t H, int32_t W,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* AGD_HIP_H */
