/*
 * emst_oracle.c -- CPU restatement of the reference EMST path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the timed CPU baseline).  The
 * product library never links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off (see oracle/Makefile).  Contraction is
 * disabled so that every fused multiply-add below is the explicit fmaf() the
 * text asks for, and nothing else is fused.
 *
 * Citations are relative to the reference repository root.
 */
#include "emst_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ A1 --
 * Bounding box as InputRW::readUniform / readTSP compute it
 * (emst_coalition/include/InputRW.h:541-559, 594-617, 488-505):
 * max updated on '>=', min on '<', starting from -/+INFINITY. */
void or_bbox(const float* pts, int64_t n, int dim, float* pmin, float* pmax)
{
    for (int k = 0; k < dim; ++k) { pmin[k] = INFINITY; pmax[k] = -INFINITY; }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < dim; ++k) {
            float c = pts[i * dim + k];
            if (c >= pmax[k]) pmax[k] = c;
            if (c < pmin[k]) pmin[k] = c;
        }
}

/* ------------------------------------------------------------------ A2 --
 * Rescale, emst_coalition/src/SolutionEMST.cu:305-379 (EMST_SQUARED_CM 0):
 *   wg = pMax[0]-pMin[0]   float subtraction, widened to double (:305-309)
 *   zero wg/hg -> 1 (:326-329)
 *   _wnd = sqrt(N*wg/hg)*2 (2D) or pow((N*wg*wg)/(hg*dg), 0.333) if dg != 0 (:330-336)
 *   unitX = (ceil(_wnd)+1)/wg (:339-342)
 *   p' = (p - pMin) * (float)unitX, PointE float arithmetic (:364, Node.h:574-593)
 * New extremes are tracked in double with '>=' / '<' (:365-379). */
double or_rescale(const float* pts, int64_t n, int dim, const float* pmin,
                  const float* pmax, float* out, double* newmin, double* newmax)
{
    double wg = (double)(pmax[0] - pmin[0]);
    double hg = (double)(pmax[1] - pmin[1]);
    double dg = 0;
    if (dim >= 3) dg = (double)(pmax[2] - pmin[2]);
    if (wg == 0) wg = 1;
    if (hg == 0) hg = 1;
    double nn = (double)n;
    double wnd = sqrt(nn * wg / hg) * 2;
    if (dim >= 3) {
        if (dg != 0) wnd = pow((nn * wg * wg) / (hg * dg), 0.333);
        else dg = 1;
    }
    double wndd = ceil(wnd) + 1;
    double unitX = wndd / wg;
    float fu = (float)unitX;
    for (int k = 0; k < 3; ++k) { newmin[k] = INFINITY; newmax[k] = -INFINITY; }
    if (dim < 3) { newmin[2] = 0; newmax[2] = 0; }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < dim; ++k) {
            float d = pts[i * dim + k] - pmin[k];
            float v = d * fu;
            out[i * dim + k] = v;
            if ((double)v >= newmax[k]) newmax[k] = v;
            if ((double)v < newmin[k]) newmin[k] = v;
        }
    return unitX;
}

/* ------------------------------------------------------------------ A3 --
 * Extents (SolutionEMST.cu:420-430): ext[k] = (int)ceil(max'-min')+1, pc = ext/2.
 * ViewGridQuadMD(pc, ext, R) integer recurrences (ViewGrid.h:251-265):
 *   PC = (pc + R-1)/R; ext_base = (PC+1) + ((ext-pc) + R-1)/R;
 *   PCD = PC/2;        ext_dual = (PCD+1) + ((ext_base-PC) + 1)/2.
 * FDual(P) = (2P + PC%2)*R - PC*R + pc (ViewGrid.h:267-293). */
static int32_t g_fdual(const or_grid* g, int k, int32_t P)
{
    return (2 * P + g->PC[k] % 2) * g->R - g->PC[k] * g->R + g->pc[k];
}

void or_grid_init(or_grid* g, int dim, int R, const double* newmin, const double* newmax)
{
    memset(g, 0, sizeof(*g));
    g->dim = dim;
    g->R = R;
    for (int k = 0; k < 3; ++k) {
        g->ext[k] = 1;
        g->ext_dual[k] = 1;
    }
    for (int k = 0; k < dim; ++k) {
        g->ext[k] = (int32_t)ceil(newmax[k] - newmin[k]) + 1;
        g->pc[k] = g->ext[k] / 2;
        g->PC[k] = (g->pc[k] + (R - 1)) / R;
        g->ext_base[k] = (g->PC[k] + 1) + ((g->ext[k] - g->pc[k]) + (R - 1)) / R;
        g->PCD[k] = g->PC[k] / 2;
        g->ext_dual[k] = (g->PCD[k] + 1) + ((g->ext_base[k] - g->PC[k]) + 1) / 2;
        g->p1[k] = (float)g_fdual(g, k, 0);
    }
    /* d_max = MAX(WD, MAX(HD, DD)), DD = 1 in 2D (EMSTOperators.h:97, ViewGrid.h:168-173) */
    int32_t dm = g->ext_dual[0];
    if (g->ext_dual[1] > dm) dm = g->ext_dual[1];
    int32_t dd = dim >= 3 ? g->ext_dual[2] : 1;
    if (dd > dm) dm = dd;
    g->d_max = dm;
    g->ncell = (int64_t)g->ext_dual[0] * g->ext_dual[1] * g->ext_dual[2];
    g->pdim = dim;
}

/* 2.5D (SolutionEMST<3,2>, emst_coalition/src/CalculateurEMST.cu:19,154-160):
 * the grid covers the first dim axes of pdim-dimensional points
 * (extents from x and y only, SolutionEMST.cu:422-430 with DimCM = 2). */
void or_grid_init_pd(or_grid* g, int dim, int pdim, int R, const double* newmin, const double* newmax)
{
    or_grid_init(g, dim, R, newmin, newmax);
    g->pdim = pdim;
}

/* findCell (ViewGrid.h:339-359): size_slab = p2[dir]-p1[dir] (= 2R);
 * li = (float)(ps[dir]-p1[dir]) + size_slab/2 in double; P = floor(li/size_slab).
 * Validity: valideAndPositiveIndex (GridOfNodes.h:314-325). */
int or_find_cell(const or_grid* g, const float* p, int32_t* cell)
{
    int ok = 1;
    for (int k = 0; k < 3; ++k) cell[k] = 0;
    for (int k = 0; k < g->dim; ++k) {
        float p2 = (float)g_fdual(g, k, 1);
        double size_slab = (double)(p2 - g->p1[k]);
        float dif = p[k] - g->p1[k];
        double li = (double)dif + size_slab / 2;
        cell[k] = (int32_t)floor(li / size_slab);
        if (cell[k] < 0 || cell[k] >= g->ext_dual[k]) ok = 0;
    }
    return ok;
}

int64_t or_cell_linear(const or_grid* g, const int32_t* c)
{
    return (int64_t)c[0] + (int64_t)g->ext_dual[0] * ((int64_t)c[1] + (int64_t)g->ext_dual[1] * c[2]);
}

static int cell_valid(const or_grid* g, const int32_t* c)
{
    for (int k = 0; k < g->dim; ++k)
        if (c[k] < 0 || c[k] >= g->ext_dual[k]) return 0;
    return 1;
}

/* ------------------------------------------------------------------ A4 --
 * d2 = (p-q)*(p-q): Point::operator- then the float dot product
 * f = 0; f += dx*dx; f += dy*dy; f += dz*dz (basic_components/include/Node.h:229-250),
 * widened to double by DistanceSquaredEuclideanP (distances_matching.h:77-86).
 * The reference GPU build uses --use_fast_math (emst_coalition/lib/lib.pro:138),
 * so each "f += a*a" after the first is the fused fmaf(a, a, f). */
float or_d2(const float* a, const float* b, int dim)
{
    float dx = a[0] - b[0];
    float f = dx * dx;
    float dy = a[1] - b[1];
    f = fmaf(dy, dy, f);
    if (dim >= 3) {
        float dz = a[2] - b[2];
        f = fmaf(dz, dz, f);
    }
    return f;
}

/* EMST_isInf (basic_components/include/macros_cuda.h:37): lexicographic
 * (d, min id, max id) with the ids already ordered by the caller. */
int or_edge_less(float d1, int32_t a1, int32_t b1, float d2, int32_t a2, int32_t b2)
{
    double x = d1, y = d2;
    if (x == y) return (a1 == a2) ? (b1 < b2) : (a1 < a2);
    return x < y;
}

void or_d2_pairs(const float* pts, int dim, const int32_t* eu, const int32_t* ev, int64_t m, float* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        out[i] = or_d2(pts + (int64_t)eu[i] * dim, pts + (int64_t)ev[i] * dim, dim);
}

static int32_t uf_find(int32_t* parent, int32_t x)
{
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

int64_t or_forest_components(int64_t n, const int32_t* eu, const int32_t* ev, int64_t m, int64_t* cycles)
{
    int64_t comps = n, cyc = 0;
    int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
    for (int64_t i = 0; i < m; ++i) {
        if (eu[i] < 0 || ev[i] < 0 || eu[i] >= n || ev[i] >= n) { cyc++; continue; }
        const int32_t a = uf_find(parent, eu[i]), b = uf_find(parent, ev[i]);
        if (a == b) { cyc++; continue; }
        parent[a < b ? b : a] = a < b ? a : b;
        comps--;
    }
    free(parent);
    if (cycles) *cycles = cyc;
    return comps;
}

static int isinf_pair(double v1, int32_t p, int32_t q1, double v2, int32_t q2)
{
    int32_t a1 = p < q1 ? p : q1, b1 = p < q1 ? q1 : p;
    int32_t a2 = p < q2 ? p : q2, b2 = p < q2 ? q2 : p;
    if (v1 == v2) return (a1 == a2) ? (b1 < b2) : (a1 < a2);
    return v1 < v2;
}

/* ------------------------------------------------------------------ A5 --
 * Cell insert: K_EMST_initializeSpiralSearch (EMSTOperators.h:80-110) puts each
 * point into cm(findCell(p)) (Cell.h:433-452).  Cells here are CSR ranges
 * (cell_start, cell_pts) with no capacity; the reference's cap of
 * MAX_CELL_SIZE = 32 (Cell.h:53) is only reported.  Points within a cell are
 * listed in index order (the reference order is atomic-order dependent and
 * does not affect results). */
int64_t or_cell_build(const or_grid* g, const float* pts, int64_t n,
                      int32_t* cell_of, int64_t* cell_start, int32_t* cell_pts)
{
    int64_t outside = 0;
    int64_t nc = g->ncell;
    memset(cell_start, 0, sizeof(int64_t) * (nc + 1));
    for (int64_t i = 0; i < n; ++i) {
        int32_t* c = cell_of + 3 * i;
        if (!or_find_cell(g, pts + i * g->pdim, c)) { outside++; continue; }
        cell_start[or_cell_linear(g, c) + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) cell_start[c + 1] += cell_start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
    memcpy(fill, cell_start, sizeof(int64_t) * nc);
    for (int64_t i = 0; i < n; ++i) {
        int32_t* c = cell_of + 3 * i;
        if (!cell_valid(g, c)) continue;
        cell_pts[fill[or_cell_linear(g, c)]++] = (int32_t)i;
    }
    free(fill);
    return outside;
}

/* ------------------------------------------------------------------ A7 --
 * Cone classification and update of one examined point q for point p,
 * NodeComputeOctantMD::computeOctants body (emst_coalition/include/NodeEMST.h:1196-1269):
 * a point goes to face (dir, sign) for every axis whose |dq| equals the max,
 * and to every slab of that face whose sign constraints (sign * dq[i] >= 0,
 * i != dir, bit order of increasing i) it meets.  The cone keeps the
 * smallest (d2, min id, max id). */
static void classify_update(int dim, int32_t p, const float* pp, int32_t q, const float* qq,
                            double v, int32_t* cid, float* cd2, int* modified)
{
    const int nslab = 1 << (dim - 1);
    float d[3];
    for (int k = 0; k < dim; ++k) d[k] = qq[k] - pp[k];
    double maxd = 0;
    for (int k = 0; k < dim; ++k)
        if (fabsf(d[k]) > maxd) maxd = fabsf(d[k]);
    for (int dir = 0; dir < dim; ++dir) {
        if (!(fabsf(d[dir]) >= maxd)) continue;
        int f = dir * 2 + ((d[dir] >= 0) ? 0 : 1);
        for (int s = 0; s < nslab; ++s) {
            int in = 1;
            unsigned mask = 1;
            for (int i = 0; i < dim; ++i) {
                if (i == dir) continue;
                double sign = (s & mask) ? -1 : 1;
                in = in && ((sign * d[i]) >= 0);
                if (!in) break;
                mask <<= 1;
            }
            if (!in) continue;
            int k = f * nslab + s;
            if (cid[k] == -1) {
                cid[k] = q; cd2[k] = (float)v; *modified = 1;
            } else if (isinf_pair(v, p, q, (double)cd2[k], cid[k])) {
                cid[k] = q; cd2[k] = (float)v; *modified = 1;
            }
        }
    }
}

/* Entry test shared by both search forms (NodeEMST.h:1051-1077): a point whose
 * stored cones hold no zero distance and at least one candidate keeps them;
 * a zero distance resets all cones and searches again (duplicates that are
 * now in the point's own component are then skipped, :1204). */
static int octants_entry(int K, int32_t* cid, float* cd2)
{
    int allinit = 1, atleastonezero = 0, nozero = 1;
    for (int k = 0; k < K; ++k) {
        if (cd2[k] == 0) { nozero = 0; atleastonezero = 1; }
        else if (cid[k] != -1) allinit = 0;
    }
    if (atleastonezero)
        for (int k = 0; k < K; ++k) { cid[k] = -1; cd2[k] = INFINITY; }
    if (nozero && !allinit) return 0;
    return 1;
}

/* Ring search, NodeEMST.h:1079-1281 (default build EMST_SPIRAL_SEARCH_KD_SLAB_2,
 * SolutionEMST.h:55-56).  Rings r = d_min(=0) .. d_max; faces in order
 * +x,-x,+y,-y,+z,-z; per face: close it when its base cell leaves the grid
 * (:1106-1115); close each slab holding a candidate when the squared distance
 * from p to the ring's face plane exceeds the candidate's d2 (:1118-1144);
 * close the face when all its slabs are closed (:1147-1157); otherwise scan the
 * (r+1)^(D-1) cells of every open slab (:1162-1276). */
int or_octants_ref_point(const or_grid* g, const float* pts, const int32_t* cell_of,
                         const int64_t* cell_start, const int32_t* cell_pts,
                         const int32_t* label, int32_t p, int32_t* cid_out, float* cd2_out,
                         int64_t* examined)
{
    const int dim = g->dim, pdim = g->pdim;
    const int nface = 2 * dim, nslab = 1 << (dim - 1), K = nface * nslab;
    int32_t cid[OR_MAXK];
    float cd2[OR_MAXK];
    memcpy(cid, cid_out, sizeof(int32_t) * K);
    memcpy(cd2, cd2_out, sizeof(float) * K);
    if (!octants_entry(K, cid, cd2)) return 0;

    const float* pp = pts + (int64_t)p * pdim;
    const int32_t* pc = cell_of + 3 * (int64_t)p;
    int faceV[6] = {0, 0, 0, 0, 0, 0};
    int slabC[6][4];
    memset(slabC, 0, sizeof(slabC));
    int modified = 0;
    int64_t ex = 0;
    for (int32_t radius = 0; radius <= g->d_max; ++radius) {
        int all = 1;
        for (int f = 0; f < nface; ++f) all = all && faceV[f];
        if (all) break;
        for (int f = 0; f < nface; ++f) {
            if (faceV[f]) continue;
            int dir = f / 2;
            int sens = (f % 2) ? -1 : 1;
            int32_t base[3] = {pc[0], pc[1], pc[2]};
            base[dir] += radius * sens;
            if (!cell_valid(g, base)) {
                faceV[f] = 1;
                for (int s = 0; s < nslab; ++s) slabC[f][s] = 1;
            }
            if (!faceV[f] && radius > 0) {
                double ppsj = pp[dir];
                float p2 = (float)g_fdual(g, dir, 1);
                double size_slab = (double)(p2 - g->p1[dir]);
                float pbase = (float)g_fdual(g, dir, base[dir]);
                double dist = fabs((double)pbase - ppsj) - size_slab / 2;
                dist *= dist;
                for (int s = 0; s < nslab; ++s)
                    if (cid[f * nslab + s] != -1 && dist > (double)cd2[f * nslab + s])
                        slabC[f][s] = 1;
            }
            if (!faceV[f]) {
                int v = 1;
                for (int s = 0; s < nslab; ++s) v = v && slabC[f][s];
                if (v) faceV[f] = 1;
            }
            if (faceV[f]) continue;
            for (int s = 0; s < nslab; ++s) {
                if (slabC[f][s]) continue;
                int32_t sign[3] = {0, 0, 0};
                int32_t ext[3] = {1, 1, 1};
                unsigned mask = 1;
                for (int i = 0; i < dim; ++i) {
                    if (i == dir) continue;
                    sign[i] = (s & mask) ? -1 : 1;
                    ext[i] = radius + 1;
                    mask <<= 1;
                }
                for (int32_t i2 = 0; i2 < ext[2]; ++i2)
                    for (int32_t i1 = 0; i1 < ext[1]; ++i1)
                        for (int32_t i0 = 0; i0 < ext[0]; ++i0) {
                            int32_t c[3] = {base[0] + i0 * sign[0], base[1] + i1 * sign[1],
                                            base[2] + i2 * sign[2]};
                            if (dim < 3) c[2] = 0;
                            if (!cell_valid(g, c)) continue;
                            int64_t lc = or_cell_linear(g, c);
                            for (int64_t j = cell_start[lc]; j < cell_start[lc + 1]; ++j) {
                                int32_t q = cell_pts[j];
                                if (q == p) continue;
                                const float* qq = pts + (int64_t)q * pdim;
                                double v = or_d2(pp, qq, pdim);
                                ex++;
                                if (v != 0 || label[q] != label[p])
                                    classify_update(dim, p, pp, q, qq, v, cid, cd2, &modified);
                            }
                        }
            }
        }
    }
    if (examined) *examined += ex;
    if (modified) {
        memcpy(cid_out, cid, sizeof(int32_t) * K);
        memcpy(cd2_out, cd2, sizeof(float) * K);
    }
    return 1;
}

/* Exact form of the same contract: every q != p (with the same duplicate
 * rule) is examined, so each cone holds its true minimum. */
int or_octants_exact_point(const float* pts, int64_t n, int dim, int cdim, const int32_t* label,
                           int32_t p, int32_t* cid_out, float* cd2_out, int64_t* examined)
{
    const int K = 2 * cdim * (1 << (cdim - 1));
    int32_t cid[OR_MAXK];
    float cd2[OR_MAXK];
    memcpy(cid, cid_out, sizeof(int32_t) * K);
    memcpy(cd2, cd2_out, sizeof(float) * K);
    if (!octants_entry(K, cid, cd2)) return 0;
    const float* pp = pts + (int64_t)p * dim;
    int modified = 0;
    for (int64_t q = 0; q < n; ++q) {
        if (q == p) continue;
        const float* qq = pts + q * dim;
        double v = or_d2(pp, qq, dim);
        if (v != 0 || label[q] != label[p])
            classify_update(cdim, p, pp, (int32_t)q, qq, v, cid, cd2, &modified);
    }
    if (examined) *examined += n - 1;
    if (modified) {
        memcpy(cid_out, cid, sizeof(int32_t) * K);
        memcpy(cd2_out, cd2, sizeof(float) * K);
    }
    return 1;
}

typedef struct {
    int32_t* cell_of;
    int64_t* cell_start;
    int32_t* cell_pts;
} or_cells;

static int64_t octants_round(int mode, const or_grid* g, const or_cells* cs, const float* pts,
                             int64_t n, int dim, int cdim, const int32_t* label, int32_t* cid,
                             float* cd2, int64_t* examined)
{
    const int K = 2 * cdim * (1 << (cdim - 1));
    int64_t searched = 0, ex = 0;
    /* every point's search reads shared state and writes only its own cones,
     * so the points run in parallel (OpenMP; one thread without it) */
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : searched, ex)
    for (int64_t p = 0; p < n; ++p) {
        int64_t e = 0;
        if (mode == 1)
            searched += or_octants_ref_point(g, pts, cs->cell_of, cs->cell_start, cs->cell_pts,
                                             label, (int32_t)p, cid + p * K, cd2 + p * K, &e);
        else
            searched += or_octants_exact_point(pts, n, dim, cdim, label, (int32_t)p, cid + p * K,
                                               cd2 + p * K, &e);
        ex += e;
    }
    if (examined) *examined += ex;
    return searched;
}

/* Threads of the parallel searches (1 without OpenMP); n > 0 sets them. */
int or_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static void cells_alloc(const or_grid* g, const float* pts, int64_t n, or_cells* cs,
                        int64_t* overfull)
{
    cs->cell_of = (int32_t*)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
    cs->cell_start = (int64_t*)malloc(sizeof(int64_t) * (g->ncell + 1));
    cs->cell_pts = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    or_cell_build(g, pts, n, cs->cell_of, cs->cell_start, cs->cell_pts);
    if (overfull) {
        *overfull = 0;
        for (int64_t c = 0; c < g->ncell; ++c)
            if (cs->cell_start[c + 1] - cs->cell_start[c] > 32) (*overfull)++;
    }
}

static void cells_free(or_cells* cs)
{
    free(cs->cell_of);
    free(cs->cell_start);
    free(cs->cell_pts);
}

void or_octants_all(int mode, const or_grid* g, const float* pts, int64_t n, int dim, int cdim,
                    const int32_t* label, int32_t* cid, float* cd2, int64_t* examined)
{
    or_cells cs = {0, 0, 0};
    if (mode == 1) cells_alloc(g, pts, n, &cs, 0);
    octants_round(mode, g, &cs, pts, n, dim, cdim, label, cid, cd2, examined);
    if (mode == 1) cells_free(&cs);
}

/* Cycle rule (an addition: the reference has no counterpart).  Each
 * component root r with a winner w (best[r]) points at the component of w's
 * target; a mutual pair (the two winners chose the same edge) is resolved by
 * the hook itself.  With cones over every axis of the distance, a
 * component's chosen edge is its shortest outgoing edge, and the choices can
 * close no other cycle.  In the 2.5D mode (xy wedges, 3D distances) they can:
 * a cycle of >= 3 components, or of 2 components with two different edges,
 * on which the reference's hook (EMSTOperators.h:1267-1303) would build a
 * cyclic label chain that its flatten never leaves.  On every such cycle the
 * component whose chosen edge is the largest under (d2, min id, max id) does
 * not hook this round: cut[r] = 1 for its root.  Every other hook is the
 * reference's, so where no such cycle forms the result is unchanged. */
static void cycle_rule(int64_t n, const int32_t* label, const int32_t* corr, const float* dist,
                       const int32_t* best, const int32_t* fixed, int32_t* tgt, int32_t* state,
                       int32_t* cut, int32_t* path)
{
    for (int64_t r = 0; r < n; ++r) {
        cut[r] = 0;
        state[r] = 0;
        tgt[r] = -1;
        int32_t w = best[r];
        if (w == -1) continue;
        int32_t q = corr[w];
        if (fixed[q] && corr[q] == w) continue;     /* mutual pair */
        tgt[r] = label[q];
    }
    for (int64_t r0 = 0; r0 < n; ++r0) {
        if (tgt[r0] == -1 || state[r0] != 0) continue;
        int64_t len = 0;
        int32_t r = (int32_t)r0;
        while (r != -1 && state[r] == 0) {
            state[r] = 1;
            path[len++] = r;
            r = tgt[r];
        }
        if (r != -1 && state[r] == 1) {
            /* the path closed on itself: the cycle runs from r to the path's end */
            int64_t k = 0;
            while (path[k] != r) ++k;
            int32_t worst = -1;
            for (int64_t i = k; i < len; ++i) {
                int32_t c = path[i], w = best[c], q = corr[w];
                if (worst == -1) { worst = c; continue; }
                int32_t v = best[worst], t = corr[v];
                if (or_edge_less(dist[v], v < t ? v : t, v < t ? t : v,
                                 dist[w], w < q ? w : q, w < q ? q : w))
                    worst = c;
            }
            cut[worst] = 1;
        }
        for (int64_t i = 0; i < len; ++i) state[path[i]] = 2;
    }
}

/* ------------------------------------------------------------ A8-A14 --
 * One Boruvka round per iteration, SolutionEMST::activate
 * (emst_coalition/src/SolutionEMSTOperators.cu:156-422):
 *   flatten (K_EMST_flatten_DST_1, EMSTOperators.h:412-433, launcher :1586-1599)
 *   termination: count labels != root(label[0]); stop at 0 or iteration == cap
 *     (EMSTOperators.h:1517-1536; SolutionEMSTOperators.cu:194-227)
 *   computeOctants + search (NodeEMST.h:1018-1384)
 *   component list + min pair (NodeEMST.h:817-955; CREATE_RESTRICTED_COMP_LST 1)
 *   hook/union (K_EMST_connectComponentAndUnion_2, EMSTOperators.h:1267-1303,
 *     parent = label snapshot, launcher :1802-1814).
 * Labels are point indices (the reference stores 4*i; same order). */
int64_t or_boruvka(int mode, const or_grid* g, const float* pts, int64_t n, int dim, int cdim,
                   int max_rounds, int32_t* eu, int32_t* ev, int32_t* eround,
                   int32_t* label_out, int32_t* cand_out, or_trace* tr)
{
    const int K = 2 * cdim * (1 << (cdim - 1));
    if (max_rounds > OR_MAX_ROUNDS) max_rounds = OR_MAX_ROUNDS;
    or_trace local;
    if (!tr) tr = &local;
    memset(tr, 0, sizeof(*tr));
    if (n <= 0) return 0;
    int32_t* label = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* corr = (int32_t*)malloc(sizeof(int32_t) * n);
    float* dist = (float*)malloc(sizeof(float) * n);
    int32_t* fixed = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* best = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* cid = (int32_t*)malloc(sizeof(int32_t) * n * K);
    float* cd2 = (float*)malloc(sizeof(float) * n * K);
    int32_t* tgt = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* state = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* cut = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* path = (int32_t*)malloc(sizeof(int32_t) * n);
    or_cells cs = {0, 0, 0};
    if (mode == 1) cells_alloc(g, pts, n, &cs, &tr->overfull_cells);
    struct timespec t_start, t_end;
    clock_gettime(CLOCK_MONOTONIC, &t_start);
    for (int64_t i = 0; i < n; ++i) {
        label[i] = (int32_t)i;              /* K_EMST_initDisjointSet, EMSTOperators.h:330-343 */
        corr[i] = -1;                        /* SolutionEMST.cu:464-466 */
        dist[i] = INFINITY;
    }
    for (int64_t i = 0; i < n * K; ++i) { cid[i] = -1; cd2[i] = INFINITY; }

    int64_t m = 0;
    int iteration = 0;
    for (;;) {
        /* flatten: chase each label to its root on a snapshot */
        for (int64_t i = 0; i < n; ++i) {
            int32_t r = (int32_t)i;
            while (r != label[r]) r = label[r];
            parent[i] = r;
        }
        memcpy(label, parent, sizeof(int32_t) * n);
        /* termination */
        int32_t r0 = label[0];
        int64_t ncomp = 0;
        for (int64_t i = 0; i < n; ++i) ncomp += (label[i] != r0);
        if (ncomp == 0 || iteration == max_rounds) break;
        iteration++;
        /* find min 1: computeOctants then search */
        int64_t ex = 0;
        tr->searched[iteration - 1] = octants_round(mode, g, &cs, pts, n, dim, cdim, label, cid, cd2, &ex);
        tr->examined[iteration - 1] = ex;
        for (int64_t p = 0; p < n; ++p) {
            int32_t pcorr = corr[p];
            if (pcorr != -1 && label[pcorr] != label[p]) continue;   /* NodeEMST.h:1315-1318 */
            corr[p] = -1;
            dist[p] = INFINITY;
            int32_t minq = -1;
            double mind = INFINITY;
            for (int k = 0; k < K; ++k) {
                int32_t q = cid[p * K + k];
                if (q == -1 || label[q] == label[p]) continue;
                double v = cd2[p * K + k];
                if (minq == -1 || isinf_pair(v, (int32_t)p, q, mind, minq)) { minq = q; mind = v; }
            }
            corr[p] = minq;
            dist[p] = (float)mind;
        }
        /* find min 2: per-component min over members with a candidate */
        for (int64_t i = 0; i < n; ++i) { fixed[i] = 0; best[i] = -1; }
        for (int64_t p = 0; p < n; ++p) {
            if (corr[p] == -1) continue;
            int32_t r = label[p];
            int32_t b = best[r];
            if (b == -1) { best[r] = (int32_t)p; continue; }
            int32_t a1 = p < corr[p] ? (int32_t)p : corr[p], b1 = p < corr[p] ? corr[p] : (int32_t)p;
            int32_t a2 = b < corr[b] ? b : corr[b], b2 = b < corr[b] ? corr[b] : b;
            if (or_edge_less(dist[p], a1, b1, dist[b], a2, b2)) best[r] = (int32_t)p;
        }
        for (int64_t r = 0; r < n; ++r)
            if (best[r] != -1) fixed[best[r]] = 1;
        cycle_rule(n, label, corr, dist, best, fixed, tgt, state, cut, path);
        /* hook / union */
        memcpy(parent, label, sizeof(int32_t) * n);
        for (int64_t p = 0; p < n; ++p) {
            if (!fixed[p]) continue;
            if (cut[parent[p]]) { tr->cycles_cut++; continue; }
            int32_t q = corr[p];
            int32_t rp = parent[p], rq = parent[q];
            if (!fixed[q] || corr[q] != p) {
                label[rp] = rq;
                eu[m] = (int32_t)p; ev[m] = q; eround[m] = iteration; m++;
            } else if (p < q) {
                label[rq] = rp;
                eu[m] = (int32_t)p; ev[m] = q; eround[m] = iteration; m++;
            }
        }
        tr->cum_edges[iteration - 1] = (int32_t)m;
    }
    tr->rounds = iteration;
    clock_gettime(CLOCK_MONOTONIC, &t_end);
    tr->ms_after_cells = (double)(t_end.tv_sec - t_start.tv_sec) * 1e3 + (double)(t_end.tv_nsec - t_start.tv_nsec) * 1e-6;
    if (label_out) memcpy(label_out, label, sizeof(int32_t) * n);
    if (cand_out) memcpy(cand_out, cid, sizeof(int32_t) * n * K);
    for (int64_t i = 0; i < m; ++i)
        if (eu[i] > ev[i]) { int32_t t = eu[i]; eu[i] = ev[i]; ev[i] = t; }
    if (mode == 1) cells_free(&cs);
    free(label); free(parent); free(corr); free(dist); free(fixed); free(best);
    free(cid); free(cd2); free(tgt); free(state); free(cut); free(path);
    return m;
}

/* ---------------------------------------------- complete-graph MSTs ---- */
int64_t or_prim_float(const float* pts, int64_t n, int dim, int32_t* eu, int32_t* ev)
{
    if (n <= 1) return 0;
    char* in = (char*)calloc(n, 1);
    float* kd = (float*)malloc(sizeof(float) * n);
    int32_t* kfrom = (int32_t*)malloc(sizeof(int32_t) * n);
    in[0] = 1;
    for (int64_t v = 1; v < n; ++v) { kd[v] = or_d2(pts, pts + v * dim, dim); kfrom[v] = 0; }
    int64_t m = 0;
    for (int64_t it = 1; it < n; ++it) {
        int64_t bv = -1;
        for (int64_t v = 0; v < n; ++v) {
            if (in[v]) continue;
            if (bv == -1) { bv = v; continue; }
            int32_t a1 = kfrom[v] < v ? kfrom[v] : (int32_t)v, b1 = kfrom[v] < v ? (int32_t)v : kfrom[v];
            int32_t a2 = kfrom[bv] < bv ? kfrom[bv] : (int32_t)bv, b2 = kfrom[bv] < bv ? (int32_t)bv : kfrom[bv];
            if (or_edge_less(kd[v], a1, b1, kd[bv], a2, b2)) bv = v;
        }
        in[bv] = 1;
        eu[m] = kfrom[bv] < bv ? kfrom[bv] : (int32_t)bv;
        ev[m] = kfrom[bv] < bv ? (int32_t)bv : kfrom[bv];
        m++;
        const float* pb = pts + bv * dim;
        for (int64_t w = 0; w < n; ++w) {
            if (in[w]) continue;
            float d = or_d2(pb, pts + w * dim, dim);
            int32_t a1 = bv < w ? (int32_t)bv : (int32_t)w, b1 = bv < w ? (int32_t)w : (int32_t)bv;
            int32_t a2 = kfrom[w] < w ? kfrom[w] : (int32_t)w, b2 = kfrom[w] < w ? (int32_t)w : kfrom[w];
            if (or_edge_less(d, a1, b1, kd[w], a2, b2)) { kd[w] = d; kfrom[w] = (int32_t)bv; }
        }
    }
    free(in); free(kd); free(kfrom);
    return m;
}

/* Kruskal: candidate keys (d2 bits << 32 | min id) sorted by a stable LSD
 * radix sort on 16-bit digits, runs of equal keys ordered by max id, then
 * union-find in key order.  d2 >= 0, so its float bits order as the value. */
int64_t or_kruskal(const float* pts, int64_t n, int dim, const int32_t* cu, const int32_t* cv, int64_t m,
                   int32_t* eu, int32_t* ev)
{
    if (n <= 1 || m <= 0) return 0;
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * m);
    uint64_t* key2 = (uint64_t*)malloc(sizeof(uint64_t) * m);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * m);
    uint32_t* idx2 = (uint32_t*)malloc(sizeof(uint32_t) * m);
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * 65537);
    int64_t mm = 0;
    for (int64_t i = 0; i < m; ++i) {
        const int32_t a = cu[i] < cv[i] ? cu[i] : cv[i], b = cu[i] < cv[i] ? cv[i] : cu[i];
        if (a == b || a < 0 || b >= n) continue;
        float d = or_d2(pts + (int64_t)a * dim, pts + (int64_t)b * dim, dim);
        uint32_t bits;
        memcpy(&bits, &d, 4);
        key[mm] = ((uint64_t)bits << 32) | (uint32_t)a;
        idx[mm] = (uint32_t)i;
        mm++;
    }
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = 16 * pass;
        memset(cnt, 0, sizeof(int64_t) * 65537);
        for (int64_t i = 0; i < mm; ++i) cnt[((key[i] >> sh) & 0xffff) + 1]++;
        for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < mm; ++i) {
            const int64_t o = cnt[(key[i] >> sh) & 0xffff]++;
            key2[o] = key[i];
            idx2[o] = idx[i];
        }
        uint64_t* tk = key; key = key2; key2 = tk;
        uint32_t* ti = idx; idx = idx2; idx2 = ti;
    }
    /* equal (d2, min id): order by max id (insertion sort within the run) */
    for (int64_t s0 = 0; s0 < mm;) {
        int64_t s1 = s0 + 1;
        while (s1 < mm && key[s1] == key[s0]) s1++;
        for (int64_t i = s0 + 1; i < s1; ++i) {
            const uint32_t x = idx[i];
            const int32_t bx = cu[x] < cv[x] ? cv[x] : cu[x];
            int64_t j = i - 1;
            while (j >= s0) {
                const uint32_t y = idx[j];
                const int32_t by = cu[y] < cv[y] ? cv[y] : cu[y];
                if (by <= bx) break;
                idx[j + 1] = y;
                j--;
            }
            idx[j + 1] = x;
        }
        s0 = s1;
    }
    int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int64_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
    int64_t e = 0;
    for (int64_t i = 0; i < mm && e < n - 1; ++i) {
        const uint32_t x = idx[i];
        const int32_t a = cu[x] < cv[x] ? cu[x] : cv[x], b = cu[x] < cv[x] ? cv[x] : cu[x];
        const int32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
        if (ra == rb) continue;
        parent[ra < rb ? rb : ra] = ra < rb ? ra : rb;
        eu[e] = a;
        ev[e] = b;
        e++;
    }
    free(parent); free(key); free(key2); free(idx); free(idx2); free(cnt);
    return e;
}

static double d2_fp64(const float* a, const float* b, int dim)
{
    double s = 0;
    for (int k = 0; k < dim; ++k) {
        double d = (double)a[k] - (double)b[k];
        s += d * d;
    }
    return s;
}

int64_t or_prim_fp64(const float* pts, int64_t n, int dim, int32_t* eu, int32_t* ev, double* weight)
{
    *weight = 0;
    if (n <= 1) return 0;
    char* in = (char*)calloc(n, 1);
    double* kd = (double*)malloc(sizeof(double) * n);
    int32_t* kfrom = (int32_t*)malloc(sizeof(int32_t) * n);
    in[0] = 1;
    for (int64_t v = 1; v < n; ++v) { kd[v] = d2_fp64(pts, pts + v * dim, dim); kfrom[v] = 0; }
    int64_t m = 0;
    double w = 0;
    for (int64_t it = 1; it < n; ++it) {
        int64_t bv = -1;
        for (int64_t v = 0; v < n; ++v)
            if (!in[v] && (bv == -1 || kd[v] < kd[bv])) bv = v;
        in[bv] = 1;
        eu[m] = kfrom[bv] < bv ? kfrom[bv] : (int32_t)bv;
        ev[m] = kfrom[bv] < bv ? (int32_t)bv : kfrom[bv];
        w += sqrt(kd[bv]);
        m++;
        const float* pb = pts + bv * dim;
        for (int64_t x = 0; x < n; ++x) {
            if (in[x]) continue;
            double d = d2_fp64(pb, pts + x * dim, dim);
            if (d < kd[x]) { kd[x] = d; kfrom[x] = (int32_t)bv; }
        }
    }
    *weight = w;
    free(in); free(kd); free(kfrom);
    return m;
}

/* ----------------------------------------------------------------- A15 --
 * Edge length sqrt(double((p-q)*(p-q))) on the original coordinates, the
 * DistanceEuclidean functor used by K_EMST_evaluate_ST
 * (EMSTOperators.h:262-292, distances_matching.h:53-62); the reference sums
 * per node then over nodes and halves (SolutionEMST.cu:566-573).  Here each
 * edge is summed once, in the order given. */
double or_weight(const float* orig, int dim, const int32_t* eu, const int32_t* ev, int64_t m)
{
    double s = 0;
    for (int64_t i = 0; i < m; ++i)
        s += sqrt((double)or_d2(orig + (int64_t)eu[i] * dim, orig + (int64_t)ev[i] * dim, dim));
    return s;
}
