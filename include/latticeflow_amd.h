/*
 * latticeflow_amd.h -- C ABI of the MI355X lattice min-cut library.
 *
 * Plain C types only (pointers, sizes, int codes).  Every entry point returns
 * LF_OK (0) or a negative LF_ERR_* code; the message of the last failure on the
 * calling thread is available from lf_last_error().
 *
 * Each declaration names the reference interface it stands in for
 * (hoytak/latticeflow; paths relative to the reference root).
 */
#ifndef LATTICEFLOW_AMD_H
#define LATTICEFLOW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define LF_OK 0
#define LF_ERR_BAD_NAME -1       /* BadSolverOrKernelName (solver/interface.hpp:152-168) */
#define LF_ERR_BAD_ARG -2        /* ValueError in pylatticeflow.pyx:157-267 */
#define LF_ERR_NONSUBMODULAR -3  /* E00 + E11 > E01 + E10 (assert in network_flow_node.hpp:450) */
#define LF_ERR_OVERFLOW -4       /* capacities do not fit the requested cap_bytes */
#define LF_ERR_DEVICE -5         /* HIP runtime failure / no GPU */
#define LF_ERR_STATE -6          /* call order (e.g. getCut before run) */
#define LF_ERR_UNSUPPORTED -7    /* kernel/solver combination without a device path */
#define LF_ERR_OUT_OF_RANGE -8   /* std::out_of_range (solver/tv/tv_solver.hpp:202-206) */

/* which cut lf_get_cut returns */
#define LF_CUT_SMIN 0 /* minimal source side: residual reachability from excess (canonical) */
#define LF_CUT_SMAX 1 /* maximal source side: nodes that cannot reach a sink */

typedef struct lf_handle lf_handle;

/* Run-time tuning of the device solver (chosen by the autotuner when zero). */
typedef struct {
  int32_t rounds_per_relabel; /* push/relabel rounds between exact global relabels */
  int32_t local_iters;        /* push/relabel iterations per tile per round, in LDS */
  int32_t max_cycles;         /* cap on global-relabel cycles (0 = none) */
  int32_t tile_shape;         /* push-tile shape: 0 = default (shape 1; shape 2 for parametric /
                                 TV networks on 2D radius-1 lattices), 1..3 = the shapes of
                                 lf_autotune; for radius-1 kernels 2D 16x32 / 16x16 / 32x32,
                                 3D 4x4x16 / 2x8x16 / 4x8x8 (4x8x32 / 2x8x32 / 4x16x16 with
                                 6 directions) */
  int32_t relabel_launches;   /* launches after which a global relabel between push rounds stops
                                 (0 = default: 2 for min-cut solves, 4 for parametric / TV
                                 networks); the relabel that decides termination always runs
                                 to its fixpoint */
} lf_tuning;

/* Solver statistics of the last lf_run. */
typedef struct {
  int64_t cycles;        /* global relabels */
  int64_t rounds;        /* push/relabel round launches */
  int64_t bfs_launches;  /* tile BFS launches over all global relabels */
  int64_t flood_launches;/* S_min flood launches */
  int64_t tiles_pushed;  /* tiles processed (not skipped) over all push launches */
  int64_t tile_nodes;    /* nodes per tile of the push kernel */
  int64_t bytes_per_node;/* algorithmic HBM bytes per processed node of one push launch */
  int64_t reserved;
  double seconds;        /* wall time of the solve (device-synchronised) */
  double push_seconds;   /* device time of all push launches (hipEvents on the solver stream) */
  double bfs_seconds;    /* wall time in global relabels (BFS + delivery + counting) */
  double flood_seconds;  /* wall time of the S_min flood */
  int64_t bfs_visits;          /* node visits of the exact global-relabel BFS (a tile's pull pass over
                                  its nodes plus every node its relaxation steps list) */
  int64_t bfs_bytes_per_visit; /* algorithmic HBM bytes per BFS visit: k residuals, k neighbour
                                  labels and the node's label (k c + 4 k + 4) */
  double bfs_device_seconds;   /* device time of the BFS launches (hipEvents around each batch) */
} lf_stats;

/* Statistics of a guided divide and conquer (lf_tv_path_guided). */
typedef struct lf_guided_stats {
  int64_t regions;          /* regions of the initial partition */
  int64_t repairs;          /* checks that found failing cut arcs (each followed by a merge) */
  int64_t failing_arcs;     /* failing cut arcs over all checks */
  int64_t layers_merged;    /* checks after which every still-failing layer was merged whole */
  int64_t fallback;         /* 1 when the partition was dropped (region table full) */
  double partition_seconds; /* building the partition and its network */
} lf_guided_stats;

/* Message of the last failure on this thread ("" when none). */
const char *lf_last_error(void);

/* ---- names and kernels: solver/interface.hpp:185-190 --------------------- */
/* isValidSolver / isValidKernel: 1 or 0. */
int lf_is_valid_solver(const char *name);
int lf_is_valid_kernel(const char *name);
/* kernelDimension (interface.hpp:188); < 0 on a bad name. */
int lf_kernel_dimension(const char *kernel);
/* validKernelNames / validSolverNames (interface.hpp:189-190): writes a
 * '\n'-separated list into buf (NUL-terminated, truncated to buflen) and
 * returns the full length needed. */
int lf_valid_kernel_names(char *buf, size_t buflen);
int lf_valid_solver_names(char *buf, size_t buflen);
/* Number of directions k of a kernel, and its offsets in generator order
 * (k * ndim ints) and geocut weights (k doubles; 1.0 when not applicable). */
int lf_kernel_size(const char *kernel);
int lf_kernel_offsets(const char *kernel, int32_t *offsets, double *geocut_weights);

/* ---- lattice handle: getLatticeFlowInterface (interface.hpp:171-183) ----- */
/* solver: "LatticeEnergyMinimizer" (binary min-cut) or "LatticeLevelReductions"
 * (parametric divide-and-conquer, pf_policies.hpp:30-46).  cap_bytes: 4 or 8,
 * the integer width of device capacities (8 = the reference default `long`). */
int lf_create(lf_handle **out, const char *solver, const char *kernel, int ndim,
              const int64_t *shape, int cap_bytes);
int lf_destroy(lf_handle *h);

/* nDimensions / nNodes / nFillingEdges / edgeDeltas (interface.hpp:24-45) */
int lf_n_dimensions(const lf_handle *h);
int64_t lf_n_nodes(const lf_handle *h);
int lf_n_filling_edges(const lf_handle *h);
/* edge_deltas[i*ndim + j], i < nFillingEdges (kernellattice.hpp:61-96 order) */
int lf_edge_deltas(const lf_handle *h, int32_t *edge_deltas);
/* all k directions in lattice order (first k/2 = edgeDeltas, then reversed) */
int lf_directions(const lf_handle *h, int32_t *offsets);

/* addEnergyPotentials(long*, long*) (interface.hpp:61, direct_interface.hpp:65-95).
 * unary[2*n + {0,1}] = E0, E1; binary[n*4*(k/2) + 4*e + {0..3}] = E00,E01,E10,E11.
 * Host pointers; accumulates like the reference (a second call adds). */
int lf_add_energy_potentials(lf_handle *h, const int64_t *unary, const int64_t *binary);
/* The same from device-resident arrays (no host copy).  `stream` (a
 * hipStream_t; NULL = the null stream) is the stream that wrote d_unary /
 * d_binary: the library waits for the work queued on it before reading them. */
int lf_add_energy_potentials_device(lf_handle *h, const int64_t *d_unary,
                                    const int64_t *d_binary, void *stream);
/* Clear all accumulated energies (a fresh lattice). */
int lf_reset_energies(lf_handle *h);

/* run() (interface.hpp:113, direct_interface.hpp:49-58).  Solves from the
 * accumulated energies; tuning may be NULL (autotuned defaults). */
int lf_set_tuning(lf_handle *h, const lf_tuning *tuning);
int lf_get_tuning(const lf_handle *h, lf_tuning *tuning);
/* Autotuner (no reference counterpart; the reference fixes its policies at
 * compile time, kernels/optimization_policies.hpp:4-41): solves the current
 * network `reps` times under each candidate tile shape and relabel schedule,
 * keeps the fastest as the handle's tuning and returns it in *best.  Every
 * candidate yields the same flow and cuts. */
int lf_autotune(lf_handle *h, int reps, lf_tuning *best);
int lf_run(lf_handle *h);
/* Invariant checks of every solve of this handle (on: nonzero): after each
 * chunk of push rounds and each exact global relabel the working state is
 * copied to the host and checked -- residuals >= 0, each arc pair's residuals
 * plus the flow in flight equal its capacity, flow conservation at every
 * node, arc masks, and exact distance labels after an exact relabel; the first
 * violation fails the solve with LF_ERR_STATE and lf_last_error names it.
 * Replaces the reference's compile-time debug verification
 * (solver/network_flow/push_relabel.hpp:132-267, _debug_forceVerifyAll);
 * for small lattices (every check copies the whole state). */
int lf_set_debug_checks(lf_handle *h, int on);
/* The same checks on host arrays in the device engine's layout (no device
 * needed): lattice (D, H, W) = dhw[0..2], k directions dirs[3*i + {0,1,2}] =
 * (dz, dy, dx) with entry i + k/2 the reverse of entry i; red0[N] and
 * alpha0[k*N] (plane-major) the base network; res[k*N], ex[N] of cap_bytes
 * each; mbox: mbox_planes planes of k*N entries of flow in flight (may be
 * NULL with 0 planes); mask[N] of mask_bytes (NULL: not checked); h[N]
 * labels, checked as exact distances when labels_exact.  LF_OK, or
 * LF_ERR_STATE with the first violation in lf_last_error. */
int lf_verify_state(const int64_t *dhw, int k, const int32_t *dirs, const int64_t *red0, const int64_t *alpha0,
                    int cap_bytes, const void *res, const void *ex, const void *mbox, int mbox_planes,
                    const void *mask, int mask_bytes, const int32_t *h, int labels_exact);
/* runTime() (interface.hpp:115): wall seconds of the last lf_run. */
int lf_run_time(const lf_handle *h, double *seconds);
int lf_get_stats(const lf_handle *h, lf_stats *stats);

/* getCut() (interface.hpp:123, direct_interface.hpp:97-106): out[n] = 1 when
 * node n (C order over the unpadded shape) is on the source ("on") side.
 * LatticeEnergyMinimizer: `which` selects S_min or S_max.
 * LatticeLevelReductions: the level-reduction cut {r < 0} (pf_solver.hpp:181-182). */
int lf_get_cut(const lf_handle *h, int32_t *out, int which);
/* The max-flow value of the last run in raw (doubled) units
 * (network_flow_node.hpp:234-240); the reference computes it but does not expose it. */
int lf_get_flow(const lf_handle *h, int64_t *flow);
/* Device pointer variant of getCut for device-resident consumers: work queued
 * on `stream` (NULL = the null stream) after this call sees the cut. */
int lf_get_cut_device(const lf_handle *h, int32_t *d_out, int which, void *stream);
/* LatticeLevelReductions: the final value r of every node (raw units) after the
 * divide and conquer, i.e. reduction + offset with all offsets reset to 0
 * (ParametricFlowSolver::run, pf_solver.hpp:33-208; PFFlowNode::r,
 * pf_flow_node.hpp:233-235).  Nodes of one level set share r up to +-1. */
int lf_get_levels(const lf_handle *h, int64_t *r_out);

/* ---- batched lattices ------------------------------------------------------ */
/* Many independent 2-D lattices of one kernel solved by one device min-cut.
 * No single call of the reference does this: a reference user creates one
 * LatticeFlowDirectInterface per lattice (getLatticeFlowInterface,
 * interface.hpp:171-183) and calls addEnergyPotentials / run / getCut on each
 * (direct_interface.hpp:49-106); a batch handle gives the same per-lattice
 * results.  shapes[2*i], shapes[2*i+1] = (n0, n1) of lattice i.  The handle
 * also takes lf_set_tuning, lf_autotune, lf_run, lf_run_time, lf_get_stats,
 * lf_reset_energies and lf_destroy; lf_n_nodes is the total node count,
 * lf_get_cut / lf_get_cut_device write every lattice's cut concatenated in
 * lattice order (each in C order), lf_get_flow is the sum of the flows. */
int lf_batch_create(lf_handle **out, const char *kernel, int n_lattices, const int64_t *shapes,
                    int cap_bytes);
/* Number of lattices of a batch handle. */
int lf_batch_size(const lf_handle *h);
/* Rows and columns of the device canvas the lattices are packed into. */
int lf_batch_canvas(const lf_handle *h, int64_t *rows_cols);
/* edgeDeltas of lattice i (its own shape fixes the order). */
int lf_batch_edge_deltas(const lf_handle *h, int i, int32_t *edge_deltas);
/* addEnergyPotentials of lattice i, in lattice i's own layout. */
int lf_batch_add_energy_potentials(lf_handle *h, int i, const int64_t *unary, const int64_t *binary);
int lf_batch_add_energy_potentials_device(lf_handle *h, int i, const int64_t *d_unary,
                                          const int64_t *d_binary, void *stream);
/* Raw max-flow value of every lattice of the last run: flows[n_lattices]. */
int lf_batch_get_flows(const lf_handle *h, int64_t *flows);

/* ---- total variation: calculate2dTV (solver/tv/tv_solver.hpp:648-681) -- */
/* Single lambda: calculate2dTV<Kernel, int64>(nx, ny, f, lambda)
 * (tv_solver.hpp:650-655 -> TVSolver::runSingleLambda :151-176).  f and R are
 * nx*ny doubles in C order; R is the reference's output (the Python layer of
 * the reference multiplies it by lambda, pylatticeflow.pyx:437-454).  The
 * kernel must be 2-D and geocut-applicable (tv_solver.hpp:51-58). */
int lf_tv_single(const char *kernel, int64_t nx, int64_t ny, const double *f, double lambda,
                 double *R, lf_stats *stats);
/* Regularization path at the given lambdas: calculate2dTV<Kernel, int64>(nx, ny,
 * f, vector<double> lambda) (tv_solver.hpp:659-681 -> TVSolver::run +
 * getRegularizationPath :181-244).  R is n_lambda * nx * ny doubles, slice l
 * for lambdas[l] (the reference's Python layer divides every slice by the last
 * lambda of the list, pylatticeflow.pyx:457-484).  All lambdas are solved
 * together on the device as layers of one stacked lattice. */
int lf_tv_path(const char *kernel, int64_t nx, int64_t ny, const double *f, int64_t n_lambda,
               const double *lambdas, double *R, lf_stats *stats);

/* The same path, solved from an initial partition of every lambda's image:
 * d_pred (device memory, n_lambda * nx * ny floats, C order per slice) holds a
 * prediction of the TV output at each lambda; neighbouring pixels whose
 * predictions differ by at most delta start in one region.  No reference
 * counterpart: the reference always starts its divide and conquer from the
 * whole lattice (ParametricFlowSolver::run, pf_solver.hpp:33-208).  Every arc
 * between two regions is taken as saturated in the order of the predictions
 * and checked once the regions are solved; regions joined by a failing arc
 * are merged and solved again (after max_repairs checks, whole lambdas), so
 * R is the same solution as lf_tv_path's for any prediction.  `stream` (NULL =
 * the null stream) is the stream that wrote d_pred: the call waits for it
 * before reading the predictions. */
int lf_tv_path_guided(const char *kernel, int64_t nx, int64_t ny, const double *f, int64_t n_lambda,
                      const double *lambdas, const float *d_pred, double delta, int max_repairs, double *R,
                      lf_stats *stats, lf_guided_stats *gstats, void *stream);

/* ---- event-driven regularization path ------------------------------------
 * The whole path of one image over lambda in [0, lambda_max], as the
 * reference's TVSolver builds it (TVSolver::run, tv_solver.hpp:114-149;
 * _constructInitialRegPathsFromSolvedLattice :300-564; TVRegPathSegment,
 * tv_regpath_node.hpp:67-1028): the exact solution at lambda_max on the
 * device, then level sets that join in closed form and split where a minimum
 * cut says so, down to lambda = 0.  Every new level set is checked by one
 * minimum cut (more where it splits): level sets of at most host_max nodes on
 * the host, larger ones on the device (0: 2048; negative: all on the
 * device).  lf_tvpath_eval then gives the path at any lambdas in
 * [0, lambda_max] (getRegularizationPath, tv_solver.hpp:181-244, traced per
 * pixel on the device): slice l for lambdas[l], negative lambdas count as 0,
 * a lambda above lambda_max is LF_ERR_OUT_OF_RANGE.  The output equals
 * lf_tv_path's with lambda_max the largest lambda of the list. */
typedef struct lf_tvpath lf_tvpath;
typedef struct lf_tvpath_stats {
  int64_t initial_regions;      /* level sets at lambda_max */
  int64_t segments;             /* level sets of the path, each alive on an interval */
  int64_t joins;                /* join events */
  int64_t splits;               /* split events */
  int64_t checks;               /* level-set checks (one minimum cut each) */
  int64_t host_checks;          /* of those, on the host */
  int64_t device_calls;         /* batched device check launches */
  int64_t dinkelbach_steps;     /* checks that located split points */
  double initial_seconds;       /* the exact solution at lambda_max */
  double device_check_seconds;  /* device checks */
  double host_check_seconds;    /* host checks */
  double seconds;               /* the whole construction */
} lf_tvpath_stats;
int lf_tvpath_create(lf_tvpath **out, const char *kernel, int64_t nx, int64_t ny, const double *f,
                     double lambda_max, int64_t host_max);
int lf_tvpath_eval(const lf_tvpath *p, int64_t n_lambda, const double *lambdas, double *R);
int lf_tvpath_get_stats(const lf_tvpath *p, lf_tvpath_stats *stats);
int lf_tvpath_destroy(lf_tvpath *p);

/* ---- learned TV predictor ---------------------------------------------------
 * No reference counterpart: a 3-layer 3x3 convolutional net (16 channels, BF16
 * matrix-core operands, FP32 accumulation) that predicts the TV output of an
 * image at a lambda, used only to build d_pred for lf_tv_path_guided.  All
 * pointers are device memory; calls are ordered on `stream`.  Inputs are n
 * images of H x W: f[n*H*W] = (image - M) / S (the TV setup's mean and scale),
 * lam[n] = log(S^2 lambda); pred / target are in the same units as f.
 * params holds lf_tvnet_param_count() floats; work is scratch memory of at
 * least lf_tvnet_work_bytes(n, H, W) bytes. */
int lf_tvnet_param_count(void);
size_t lf_tvnet_work_bytes(int n, int64_t H, int64_t W);
int lf_tvnet_forward(const float *params, int n, int64_t H, int64_t W, const float *f, const float *lam,
                     float *pred, void *work, size_t work_size, void *stream);
/* The forward pass of lf_tvnet_forward without any allocation and without a
 * workspace argument.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Its scratch (bf16 weights and
 * activations) lives in a 64 MiB static device array that every GPU the
 * library runs on holds until the process ends; n * H * W may be at most
 * LF_TVNET_NOALLOC_MAX_PIXELS (larger inputs fail with LF_ERR_BAD_ARG).  All
 * calls on one GPU share that array: calls that overlap on one GPU race.
 * Output equals lf_tvnet_forward's. */
#define LF_TVNET_NOALLOC_MAX_PIXELS 655360
int lf_tvnet_forward_noalloc(const float *params, int n, int64_t H, int64_t W, const float *f, const float *lam,
                             float *pred, void *stream);
/* Gradient of the mean squared error against target: grad[param_count],
 * loss[1] (both overwritten). */
int lf_tvnet_gradients(const float *params, int n, int64_t H, int64_t W, const float *f, const float *lam,
                       const float *target, float *grad, float *loss, void *work, size_t work_size,
                       void *stream);
/* One Adam update (beta1 0.9, beta2 0.999, eps 1e-8) of params from grad;
 * m and v are the moment buffers (zero before step 1). */
int lf_tvnet_adam(float *params, const float *grad, float *m, float *v, int step, float lr, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LATTICEFLOW_AMD_H */
