/*
 * emst.h -- C ABI of the MI355X (gfx950) EMST engine.
 *
 * Drop-in boundary for the reference's EMST entry surface: a point array goes
 * in; the spanning forest after any Boruvka round, or the finished tree, and
 * its total weight come out.  Every entry point names the reference interface
 * it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *   - Plain C types only; every call returns an emst_status (0 = EMST_OK).
 *   - A context owns its device buffers and one HIP stream; contexts are
 *     independent and may be used from different host threads (one thread
 *     per context at a time).  No process globals, nothing ever exits.
 *   - Point ids are input indices 0..n-1.  The reference stores byte offsets
 *     (i * sizeof(node)), an order-preserving map of the same ids
 *     (basic_components/include/GridOfNodes.h:454-465).
 *   - Edge order is canonical: u < v, sorted by (u, v).
 *   - Ties between squared distances break on (d2, min id, max id)
 *     (basic_components/include/macros_cuda.h:37).
 */
#ifndef EMST_AMD_H
#define EMST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EMST_OK = 0,
    EMST_E_ARG = 1,         /* bad argument (NULL, dim, n < 0, size too small) */
    EMST_E_HIP = 2,         /* HIP runtime or launch error */
    EMST_E_OOM = 3,         /* device allocation failed */
    EMST_E_NONFINITE = 4,   /* NaN or Inf coordinate */
    EMST_E_STATE = 5,       /* call out of order (e.g. step before set_points) */
    EMST_E_NODEVICE = 6,    /* no HIP device */
    EMST_E_LIMIT = 7,       /* input beyond a supported limit (n >= 2^31) */
    EMST_E_IO = 8           /* file read/write failure */
} emst_status;

/* Options.  Replaces the INI keys the EMST path reads through
 * ConfigParamsCF (coalition_framework/src/config/ConfigParamsCF.cpp:131-144,
 * basic_components/include/ConfigParams.h:313-321) and the compile-time
 * switches (emst_coalition/src/SolutionEMSTOperators.cu:17-18,
 * emst_coalition/include/SolutionEMST.h:55-56). */
typedef struct {
    int32_t dim;            /* point dimension: 2 or 3 */
    int32_t level_radius;   /* R of the reference cell grid (levelRadius); cells are 2R wide */
    int32_t max_rounds;     /* Boruvka round cap; the reference stops at 100 */
    int32_t bucket_size;    /* points per leaf cell of the search hierarchy: a power of two in
                               [8, 64], rounded up (0 = 16 for the 2D cone search, else 64) */
    int32_t trace;          /* 1: record per-round forest size, length and phase times */
    int32_t device;         /* HIP device ordinal */
    int32_t radius_mode;    /* 0 plain; 1 learned seed radius (needs emst_set_radius_model) */
    float seed_scale;       /* plain-form first search radius, in mean spacings of the point's
                               leaf cell (0 = default: 3.0, 1.0 for the 3D component search);
                               changes work, never results */
    int32_t cone_dim;       /* dimension of the cones (DimCM): 0 = the point dimension; 2 with
                               3D points is the 2.5D mode, SolutionEMST<3,2>
                               (emst_coalition/src/CalculateurEMST.cu:19,154-160,249-258):
                               wedges over the xy plane, 3D distances, the 3D rescale.  There a
                               component's chosen edges can close a cycle of components, on
                               which the reference's flatten does not end; here the component
                               whose edge is the cycle's largest under (d2, min id, max id)
                               skips that round's hook (emst_get_search_stats entry 7 counts
                               these).  Not with the learned seed radius. */
    int32_t search_split;   /* round-1 cone search launches (changes work, never results):
                               0 auto (3D cones and >= 2^21 points: 3 when at most 10% of the
                               queries are left with open cones after the first pass, else 2;
                               otherwise 1); 1 one launch for
                               every pass; 2 first pass, then the open queries resumed in place;
                               3 first pass, then the open queries compacted into full waves */
    int32_t round_search;   /* 0: round-1 cone search, candidates reused by every round (the
                               reference's default build, NodeEMST.h:960-1386); 1: per-round
                               component search, each point searching for its nearest point in
                               another component whenever its target has joined its own
                               (NodeSpiralSearchMD, NodeEMST.h:67-509, selected in the reference by
                               EMST_SPIRAL_SEARCH_KD_SLAB_2 0, SolutionEMST.h:55-56).  Both build
                               the same forests.  1 has no cone candidates and no 2.5D mode; it
                               keeps each point's nearest neighbours from the build (10 in 3D, 121
                               bytes per point; 6 in 2D, 73 bytes), which answer the early rounds
                               without a search; its
                               learned seed predicts each component's edge length per round */
    int32_t log_samples;    /* 1 (round_search 1): log the learned seed's training pairs during
                               runs (emst_get_radius_samples) */
    int32_t one_launch;     /* 1 (round_search 1, untraced, no seed model or sample log): emst_run
                               issues the whole run -- bounding boxes, rescale, sort, hierarchy,
                               neighbour lists and the first 16 Boruvka rounds -- as one HIP graph
                               launch with one host synchronisation at its end (captured on first
                               use, replayed while the point layout is the same); rounds past those
                               continue as usual.  The results are those of the default run.  For
                               batches of many small clouds and for callers that replay the same
                               shapes; not for emst_step. */
} emst_options;

/* One row of the per-round trace: the columns of the reference's stats file
 * (emst_coalition/src/SolutionEMSTRW.cu:100-165, TspResultInfo in
 * emst_coalition/include/EMSTOperators.h:54-70).  Times in milliseconds. */
typedef struct {
    int32_t iteration;
    int32_t size;               /* forest edges after this round */
    double length;              /* forest weight after this round (FP64, input coords) */
    float time_flatten;
    float time_termination;
    float time_find_next_closest;   /* cone search + per-point candidate min */
    float time_find_min_pair;       /* per-component minimum */
    float time_connect_union;       /* hooking */
    int64_t searched;               /* points whose cone search ran this round; with
                                       round_search 1, the queries of the round's
                                       component search (first and resumed dispatch) */
} emst_round_trace;

typedef struct emst_ctx emst_ctx;

/* Fills defaults: dim 3, level_radius 1, max_rounds 100, bucket_size 0 (auto),
 * trace 0, device 0, radius_mode 0. */
void emst_default_options(emst_options* opt);

/* Replaces CalculateurEMST::initialize's static set-up and the device query
 * of SolutionEMST::initialize() (emst_coalition/src/CalculateurEMST.cu:113-230,
 * emst_coalition/src/SolutionEMST.cu:17-270).  Returns NULL on failure; the
 * reason is then given by emst_last_error(NULL). */
emst_ctx* emst_create(const emst_options* opt);
int emst_destroy(emst_ctx* ctx);

/* Text of the last error on this context (or of the last failed
 * emst_create when ctx is NULL).  Never NULL. */
const char* emst_last_error(const emst_ctx* ctx);

/* Replaces SolutionEMST<DimP,DimCM>::initialize(NetLink&, PointEuclidean& pMin,
 * PointEuclidean& pMax) (emst_coalition/src/SolutionEMST.cu:273-509): takes an
 * in-memory point array (AoS, n x dim float32, like the reference's
 * adaptiveMap of PointE<D>), computes the bounding box (InputRW.h:541-559),
 * rescales (SolutionEMST.cu:305-379) and builds the search hierarchy.
 * xyz_on_device = 1: xyz is a device pointer on the context's device and the
 * context only reads it (it must stay valid until the next set_points or
 * destroy).  Resets the forest to n singletons. */
int emst_set_points(emst_ctx* ctx, const float* xyz, int64_t n, int32_t dim,
                    int32_t xyz_on_device);

/* Batched form of emst_set_points (the batched entry of the drop-in surface;
 * no counterpart in the reference, which runs one cloud per process):
 * ncl independent clouds of one dimension, packed back to back in xyz (cloud
 * c's npts[c] points follow cloud c-1's, AoS).  Global point ids run over
 * the packed array (cloud c's point j has id offset_c + j, offset_c = sum of
 * npts[<c]); every cloud is rescaled, searched and hooked on its own, so its
 * forest, weight and per-round edges equal a run of that cloud alone.
 * emst_run / emst_step advance all clouds together; emst_get_edges returns
 * global ids; emst_get_cloud_results the per-cloud figures.  The learned
 * seed radius and the reference cell grid are single-cloud only. */
int emst_set_clouds(emst_ctx* ctx, int32_t ncl, const int64_t* npts, const float* xyz, int32_t dim,
                    int32_t xyz_on_device);

/* Per cloud of the last set_clouds / set_points: FP64 forest weight (input
 * coordinates), forest edge count, and the last round that added one of its
 * edges (its completion round).  Arrays of ncl entries; any may be NULL. */
int emst_get_cloud_results(emst_ctx* ctx, double* weight, int32_t* edges, int32_t* last_round);

/* Autotuner (no counterpart in the reference, whose cell size is the fixed
 * level radius): times the build and cone search of the context's current
 * points on the device for every leaf size (points per leaf cell: 16, 32,
 * 64) and plain seed scale (2, 3, 4 leaf spacings), then every search_split
 * form (1, 2, 3) for the fastest pair, then whole runs with round_search 0
 * and with 1 at 64-point leaves (not with a learned model set, nor in the
 * 2.5D mode, where the
 * caller's choice stays); keeps the fastest choice in the context (used by
 * every later set_points / set_clouds / run) and returns the pair; best_ms =
 * its build + cone search time.  The result of a run never depends on any of
 * them; any output may be NULL. */
int emst_autotune(emst_ctx* ctx, int32_t* best_bucket, float* best_scale, float* best_ms);

/* Replaces SolutionEMST::run() (emst_coalition/src/SolutionEMSTOperators.cu:146-153):
 * Boruvka rounds until one tree remains, no component has an outgoing
 * candidate (a forest), or max_rounds is reached. */
int emst_run(emst_ctx* ctx);

/* Replaces SolutionEMST::activate() / CalculateurEMST::activate()
 * (emst_coalition/src/SolutionEMSTOperators.cu:156-422,
 * emst_coalition/src/CalculateurEMST.cu:330-429): one Boruvka round.
 * *finished = 1 when the termination test holds before the round
 * (the round is then not run). */
int emst_step(emst_ctx* ctx, int32_t* finished);

/* Replaces getLinks()/getLinks3() (emst_coalition/include/CalculateurEMST.h:52-58,
 * CalculateurEMST.cu:97-111), which return the per-node adjacency: here the
 * forest's edge list.  Pass NULL arrays to query *m only; otherwise *m is the
 * capacity on entry and the count on return.  d2 = float32 squared distance
 * in rescaled coordinates (the ordering key).  Any array may be NULL. */
int emst_get_edges(emst_ctx* ctx, int32_t* u, int32_t* v, float* d2, int64_t* m);

/* Replaces SolutionEMST::evaluate() (emst_coalition/src/SolutionEMST.cu:560-581,
 * K_EMST_evaluate_ST emst_coalition/include/EMSTOperators.h:262-292): FP64 sum
 * of sqrt(float d2) over the forest edges, on the input coordinates. */
int emst_get_weight(emst_ctx* ctx, double* weight, int64_t* size);

/* No counterpart in the reference: the entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The FP64 weight that emst_get_weight
 * computes, for an edge list the caller holds on the device: the sum over
 * i < m with eu[i] >= 0 of sqrt(double(d2)), d2 the float32 squared distance
 * of xyz[eu[i]] and xyz[ev[i]] (xyz: n x dim float32, AoS, input
 * coordinates).  xyz, eu and ev are device pointers on `device`; the work
 * runs on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns once *weight (host memory) is written.  It calls no hipMalloc or
 * hipFree: its partial sums (512 doubles, 4 KiB) live in a static device
 * array of the library, one per GPU, held until the process ends; every call
 * on a GPU overwrites all of them, and calls that overlap on one GPU race.
 * m must be below 2^31 and dim 2 or 3; on failure emst_last_error(NULL)
 * gives the reason. */
int emst_weight_noalloc(const float* xyz, int32_t dim, const int32_t* eu, const int32_t* ev, int64_t m,
                        int32_t device, void* stream, double* weight);

/* The disjoint-set labels (disjointSetMap, written by NeuralNetEMST::writeLinks
 * emst_coalition/include/NeuralNetEMST.h:310-318): root point index per point,
 * after the last round's flatten. */
int emst_get_labels(emst_ctx* ctx, int32_t* labels);

/* Replaces getAdaptiveMap()/getAdaptiveMap3() (CalculateurEMST.h:45-51): the
 * rescaled coordinates, n x dim float32, input order; also unitX. */
int emst_get_rescaled(emst_ctx* ctx, float* xyz, double* unit_x);

/* Per-round trace rows (see emst_round_trace).  *nrows: capacity in, count out. */
int emst_get_trace(emst_ctx* ctx, emst_round_trace* rows, int32_t* nrows);

/* Stage output for parity: the cone candidates of the last cone search,
 * n x K input ids (-1 = empty cone), K = 8 (2D) or 24 (3D), and their float32
 * d2 (inf = empty).  Mirrors NodeComputeOctantMD::slabP/slabDP
 * (emst_coalition/include/NodeEMST.h:960-1011). */
int emst_get_candidates(emst_ctx* ctx, int32_t* cand, float* cand_d2, int32_t* k);

/* The reference's cellular matrix for the current points: dual-cell extents
 * and, per point, its dual cell (findCell, basic_components/include/ViewGrid.h:339-359)
 * and the CSR listing of each cell (the cell insert of K_EMST_initializeSpiralSearch,
 * emst_coalition/include/EMSTOperators.h:80-110, without the 32-point cap of
 * basic_components/include/Cell.h:53).  ext_dual[3] out; cell_of n x 3 out
 * (-1 when outside); cell_start (ncell+1) and cell_pts (n) out, ncell =
 * ext_dual[0]*ext_dual[1]*ext_dual[2]; pass NULL cell_start/cell_pts to get
 * only ext_dual.  Points of a cell are listed in increasing id. */
int emst_get_cell_grid(emst_ctx* ctx, int32_t level_radius, int32_t* ext_dual,
                       int32_t* cell_of, int64_t* cell_start, int32_t* cell_pts);

/* Kernel times of the last emst_run (ms, device events on the context's
 * stream): [0] build (bbox, rescale, sort, hierarchy, coincident groups, the
 * learned seed's features and inference when set, and the candidates ranked
 * for the rounds), [1] round-1 cone search, [2] Boruvka rounds, [3] total,
 * [4] the per-round component search's search dispatches (k_comp_search)
 * summed over the rounds, [5] the component search's neighbour lists at
 * build (k_knn_first); [4] and [5] in traced runs (options.trace; 0
 * otherwise).
 * n_times: capacity in, count out. */
int emst_get_times(emst_ctx* ctx, float* times, int32_t* n_times);

/* Device time from the start of the last one_launch run (options.one_launch)
 * to the end of each of its captured rounds, in milliseconds: round r's entry
 * is ms[r - 1], so a cloud that completed in round r (emst_get_cloud_results)
 * had its forest that long after the run began.  Other runs leave no entries.
 * ms NULL: *n = the count; otherwise *n is the capacity in, count out.  No
 * counterpart in the reference. */
int emst_get_round_ends(emst_ctx* ctx, float* ms, int32_t* n);

/* Counters of the last full cone search (the work behind its bytes):
 * [0] packets (64-point waves), [1] hierarchy nodes tested, [2] leaf cells
 * entered, [3] leaf points read (summed over packets), [4] point pairs
 * evaluated (leaf points x lanes that needed the leaf), [5] points sharing
 * their location with another point, [6] 1 when those were searched as
 * groups (one search per location), 0 when every point searched itself,
 * [7] hooks withheld by the 2.5D mode's cycle rule over the rounds run so far,
 * [8] queries still open after the first pass of a split search (see
 * emst_options.search_split; 0 when the search ran as one launch),
 * [9] queries over all the per-round component search's dispatches of the
 * run so far (round_search 1; a query resumed after its first pass counts
 * twice), [10] hierarchy nodes those dispatches tested (one per node a
 * 64-query packet examined), [11] leaf cells they entered, [12] point pairs
 * they evaluated (leaf points x queries that needed the leaf), [13]-[15] the
 * same three for the neighbour lists of the component search's build
 * (pairs there: leaf points examined x 64 lanes), [16] passes of the
 * component search's packets that walked the hierarchy per query (an
 * unbounded pass, or a gathering of candidate leaves that outgrew its lists),
 * [17] packets of the neighbour lists' build that did; [10]-[17] are counted
 * in traced runs (options.trace) only, 0 otherwise.
 * No counterpart in the reference (its per-phase timers are in
 * SolutionEMSTOperators.cu:166-373).  n_stats: capacity in, count out. */
int emst_get_search_stats(emst_ctx* ctx, int64_t* stats, int32_t* n_stats);

/* ---- learned seed radius (no counterpart in the reference) -----------------
 * Cone search (round_search 0):
 * a multilayer perceptron 8 -> 64 -> 64 -> 1 (ReLU; BF16 weights and
 * activations, FP32 accumulation on the matrix cores) predicts, per point,
 * log(R / s0): R = radius within which all the point's non-empty cones have
 * their nearest point, s0 = mean point spacing of its leaf cell.  The cone
 * search starts its iterative deepening at s0 * exp(prediction + offset); the
 * result is the same for any seed, only the work changes.
 * Model: EMST_MLP_PARAMS floats, W1[8][64] | b1[64] | W2[64][64] | b2[64] |
 * w3[64] | b3 | feature mean[8] | feature 1/std[8] | seed offset.
 * Features (per point, all independent of the cloud's scale): log(d / s0),
 * d = distance to the nearest of the 8 nearest sorted (Morton-order)
 * positions; log(s1/s0), log(s2/s1), log(s3/s2) for the 1st..3rd enclosing
 * cells; log aspect of the leaf box; distance to the leaf box boundary / s0;
 * log(median distance to those 8 neighbours / s0); length of the sum of the 8
 * neighbour offsets / sum of their lengths.  Normalised features are clamped
 * to [-4, 4] before the first layer. */
#define EMST_MLP_PARAMS 4818
/* Component search (round_search 1; where the reference's ring search starts
 * from its remembered d_min, NodeEMST.h:209-214): the same network predicts,
 * every round for every searching point, log(min(D, d) / s0), D = the length
 * of its component's shortest outgoing edge, d = the distance of the point's
 * nearest point outside the component.  The point's first pass searches to
 * max(s0 * exp(prediction + offset), its lower bound); the component's edge
 * found by then bounds its other members' searches.  Features: log round;
 * log component size; log(lower bound / s0) (the previous target's
 * distance); log(extent of the component's box / s0); log(the component's
 * mean spacing (box volume / size)^(1/D) / s0); log(the component's kept
 * minimum / s0) (10^3 when none); log(s1/s0); log(s3/s0).  Lengths below
 * 10^-3 s0 count as 10^-3 s0. */

/* Use `model` for the seed radius of this context's searches (copied): a
 * cone-search model with round_search 0, a component-search model with
 * round_search 1.  NULL returns to the plain seed (seed_scale x leaf spacing;
 * with round_search 1 at least twice the lower bound). */
int emst_set_radius_model(emst_ctx* ctx, const float* model);

/* Training pairs.  round_search 0: features (m x 8) and targets (m) of every
 * point of the context's current points, from its last cone search.
 * round_search 1: the pairs logged by the last run (options.log_samples = 1:
 * every searching point of every round, subsampled to at most 2^16 per round
 * and 2^20 in all).  Pass NULL arrays to query *m; otherwise *m is the
 * capacity in, count out. */
int emst_get_radius_samples(emst_ctx* ctx, float* feat, float* target, int64_t* m);

/* Trains a model on the GPU with the hand-written forward/backward kernels:
 * mean squared error on targets, Adam (lr, betas 0.9/0.999), `epochs` passes
 * over m samples in shuffled minibatches of `batch` (a multiple of 64), He
 * initialisation from `seed`, feature normalisation from the data.  model:
 * EMST_MLP_PARAMS floats out; loss_per_epoch (optional): mean squared error
 * of each epoch. */
int emst_mlp_train(int32_t device, const float* feat, const float* target, int64_t m, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float* model, float* loss_per_epoch);

/* Forward pass on the GPU: out[i] = prediction for feat[i] (without offset). */
int emst_mlp_eval(int32_t device, const float* model, const float* feat, int64_t m, float* out);

/* Gradient of sum_i (prediction_i - target_i)^2 with respect to the model's
 * trainable parameters (first 4801 floats; the rest are returned as 0), and
 * that sum, from the GPU kernels. */
int emst_mlp_grad(int32_t device, const float* model, const float* feat, const float* target, int64_t m,
                  float* grad, double* loss);

/* Points and dimension of the context's current point set. */
int emst_get_shape(emst_ctx* ctx, int64_t* n, int32_t* dim);

/* The context's current options, including what emst_autotune chose
 * (bucket_size, seed_scale, search_split), e.g. to create further contexts
 * that run the same way.  No counterpart in the reference. */
int emst_get_options(emst_ctx* ctx, emst_options* opt);

/* ---- on-disk formats (host only; no GPU needed for the reader) ------------
 * Replaces InputRW::readFile (emst_coalition/include/InputRW.h:44-623): the
 * file name (after the last '/') is cut at its first '.' or '_' (getPos); the
 * rest picks the reader,
 * which opens base + its suffix: ".points" ("Width = W Height = H", then W*H
 * points), ".3d" / ".2d" (first token N, then N points), anything else
 * TSPLIB (NAME on line 1, "DIMENSION : N" on line 4, "index x y" from line
 * 7, up to N cities or an earlier "EOF").  dim: the solution's point dimension (2D data read by a 3D solution get
 * z = 0; a .3d file needs dim 3).  xyz NULL: *n = point count only;
 * otherwise *n is the capacity in, count out.  EMST_E_IO on a missing or
 * malformed file. */
int emst_read_points(const char* path, int32_t dim, float* xyz, int64_t* n);

/* Replaces SolutionEMST::writeSolution / NeuralNetEMST::writeLinks
 * (emst_coalition/src/SolutionEMSTRW.cu:38-62, NeuralNetEMST.h:277-326,
 * BufferLink.h:266-279): writes base.links (every node with its forest
 * neighbours, ascending), base.correspondenceMap ((-1, -1) per point, the
 * host copy the reference writes), base.disjointSetMap (4 x root index per
 * point) and base.points (rescaled coordinates as 3D points), base = path cut
 * at its first '.' or '_'. */
int emst_write_solution(emst_ctx* ctx, const char* path);

/* Synthetic clouds, identical on every host (no counterpart in the reference;
 * used by the benchmark and tests).  SplitMix64-seeded xoshiro128+,
 * u = (x >> 8) * 2^-24.  kind: 0 uniform [0,1)^dim; 1 Gaussian mixture
 * (64 centres, sigma log-uniform 0.002-0.02, 10% uniform background, clamped
 * to the unit cube); 2 power-law (16 centres, radial density r^-1.5,
 * r in [1e-4, 0.5]); 3 eight Gaussians (sigma as kind 1, no background).
 * out: n x dim float32 host memory. */
int emst_synth(int32_t kind, int64_t n, int32_t dim, uint64_t seed, float* out);

/* Library version string. */
const char* emst_version(void);

#ifdef __cplusplus
}
#endif
#endif /* EMST_AMD_H */
