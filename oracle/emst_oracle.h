/*
 * emst_oracle.h -- CPU restatement of the reference EMST path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Every function states the reference lines it restates.  Paths are relative
 * to the reference repository root.
 *
 * Point ids: the reference identifies points by byte offsets into its grids
 * (GridOfNodes.h:454-465), which are i*sizeof(Node): an order-preserving map
 * of the point index i.  The oracle uses i directly; file writers multiply by
 * the node size where the reference's files show offsets.
 */
#ifndef EMST_ORACLE_H
#define EMST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXD 3
#define OR_MAXK 24          /* cones per point: 2D faces x 2^(D-1) slabs (D = cone dimension) */
#define OR_MAX_ROUNDS 128

/* Cellular-matrix geometry: ViewGridQuadMD (basic_components/include/ViewGrid.h:251-265)
 * built on extents ceil(max'-min')+1 (emst_coalition/src/SolutionEMST.cu:420-430). */
typedef struct {
    int dim;            /* DimCM: grid and cone dimension */
    int R;              /* levelRadius */
    int32_t ext[3], pc[3], PC[3], ext_base[3], PCD[3], ext_dual[3];
    float p1[3];        /* FEuclid(FDual(0)) */
    int32_t d_max;      /* max dual extent: EMSTOperators.h:97 */
    int64_t ncell;
    int pdim;           /* DimP: point dimension (stride of the point arrays; = dim
                           except in the 2.5D mode SolutionEMST<3,2>) */
} or_grid;

typedef struct {
    int32_t rounds;                         /* Boruvka rounds executed */
    int32_t cum_edges[OR_MAX_ROUNDS];       /* forest size after round r (index r-1) */
    int64_t searched[OR_MAX_ROUNDS];        /* points whose cone search ran in round r */
    int64_t examined[OR_MAX_ROUNDS];        /* candidate points examined in round r */
    int64_t overfull_cells;                 /* cells above the reference cap of 32 */
    int64_t cycles_cut;                     /* hooks withheld by the cycle rule */
    double ms_after_cells;                  /* wall time from after the cell build to the
                                               finished forest (the span of the reference's
                                               cuda_duree, SolutionEMSTRW.cu:85-97,133-165) */
} or_trace;

/* A1 */
void or_bbox(const float* pts, int64_t n, int dim, float* pmin, float* pmax);
/* A2: returns unitX; out = rescaled coordinates; newmin/newmax = their extremes */
double or_rescale(const float* pts, int64_t n, int dim, const float* pmin,
                  const float* pmax, float* out, double* newmin, double* newmax);
/* A3: grid of dimension dim (DimCM) over points of dimension pdim (DimP) */
void or_grid_init(or_grid* g, int dim, int R, const double* newmin, const double* newmax);
void or_grid_init_pd(or_grid* g, int dim, int pdim, int R, const double* newmin, const double* newmax);
int or_find_cell(const or_grid* g, const float* p, int32_t* cell);
int64_t or_cell_linear(const or_grid* g, const int32_t* cell);
/* A4 */
float or_d2(const float* a, const float* b, int dim);
int or_edge_less(float d1, int32_t a1, int32_t b1, float d2, int32_t a2, int32_t b2);
/* A4 over an edge list: out[i] = or_d2(pts + eu[i]*dim, pts + ev[i]*dim, dim). */
void or_d2_pairs(const float* pts, int dim, const int32_t* eu, const int32_t* ev, int64_t m, float* out);
/* Union-find over the edge list: returns the number of components of the
 * graph on n vertices; *cycles = edges whose ends were already connected. */
int64_t or_forest_components(int64_t n, const int32_t* eu, const int32_t* ev, int64_t m, int64_t* cycles);
/* A5: returns number of points outside the grid (should be 0) */
int64_t or_cell_build(const or_grid* g, const float* pts, int64_t n,
                      int32_t* cell_of, int64_t* cell_start, int32_t* cell_pts);
/* A7: one point's computeOctants; returns 1 if the ring search ran */
int or_octants_ref_point(const or_grid* g, const float* pts, const int32_t* cell_of,
                         const int64_t* cell_start, const int32_t* cell_pts,
                         const int32_t* label, int32_t p, int32_t* cid, float* cd2,
                         int64_t* examined);
/* Exact cone-nearest (Yao) restatement of the same per-point contract.
 * dim = point dimension (distances), cdim = cone dimension (2 with dim 3 is
 * the 2.5D mode: wedges over the xy plane, 3D distances). */
int or_octants_exact_point(const float* pts, int64_t n, int dim, int cdim, const int32_t* label,
                           int32_t p, int32_t* cid, float* cd2, int64_t* examined);
/* Stage A7 over all points (mode 0 = exact, 1 = reference ring search). */
void or_octants_all(int mode, const or_grid* g, const float* pts, int64_t n, int dim, int cdim,
                    const int32_t* label, int32_t* cid, float* cd2, int64_t* examined);

/* A8-A14: Boruvka driver.  mode 0 = exact cone search, 1 = reference ring
 * search on grid g (g->dim = cdim, g->pdim = dim).  eu/ev/eround need n-1
 * entries.  Returns edge count; tr->cycles_cut counts hooks withheld by the
 * cycle rule (2.5D only; see or_boruvka). */
int64_t or_boruvka(int mode, const or_grid* g, const float* pts, int64_t n, int dim, int cdim,
                   int max_rounds, int32_t* eu, int32_t* ev, int32_t* eround,
                   int32_t* label_out, int32_t* cand_out, or_trace* tr);

/* Complete-graph MSTs.  Prim under the float total order (d2 float, min id,
 * max id); FP64 Prim on double distances.  Return edge count. */
int64_t or_prim_float(const float* pts, int64_t n, int dim, int32_t* eu, int32_t* ev);
int64_t or_prim_fp64(const float* pts, int64_t n, int dim, int32_t* eu, int32_t* ev,
                     double* weight);

/* Kruskal under the same total order over a candidate edge list (either
 * orientation, duplicates and self pairs allowed): writes the chosen edges
 * (u < v) in key order and returns their count.  With every pair as a
 * candidate it is the complete-graph MST (the tree of or_prim_float); with a
 * candidate set that holds a spanning tree T it returns T exactly when no
 * candidate edge could replace an edge of T (the cycle property over the
 * candidates).  Ids must be below 2^31. */
int64_t or_kruskal(const float* pts, int64_t n, int dim, const int32_t* cu, const int32_t* cv, int64_t m,
                   int32_t* eu, int32_t* ev);

/* A15 */
double or_weight(const float* orig, int dim, const int32_t* eu, const int32_t* ev, int64_t m);

/* Threads of the per-point cone searches (OpenMP); n > 0 sets them.  Returns
 * the count now in use. */
int or_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
