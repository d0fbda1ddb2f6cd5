// emst_api.hip -- host side of the C ABI (include/emst.h).
//
// A context holds one point set on one device and one HIP stream.  The build
// (bbox, rescale, Morton sort, hierarchy, round-1 cone search) runs inside
// the first emst_run/emst_step, so that a timed run starts from resident
// input points, like the reference's cuda_duree span plus its cell build.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emst.h"
#include "emst_internal.h"

using namespace emst;

namespace {

thread_local std::string g_create_error;

constexpr int kWeightBlocks = 512;

// Device buffer owned by a context (freed with it).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

// Partial sums of emst_weight_noalloc: one k_weight block each.  A static
// device array, so that entry point allocates nothing; each GPU the library
// runs on holds its own copy, at its own address, until the process ends.
// Host code never passes `scratch` itself to a kernel: it uses the address
// that hipGetSymbolAddress gives for the current device (scratch_on).
static __device__ __align__(256) double scratch[kWeightBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
hipError_t scratch_on(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace

struct emst_ctx {
    emst_options opt{};
    hipStream_t stream = nullptr;
    std::string err;
    // point set: ncl clouds (one for emst_set_points), n points in all
    int64_t n = 0;
    int dim = 0;
    int cdim = 0;                     // cone dimension: dim, or 2 over 3D points (2.5D mode)
    int K = 0;
    int bucket = 32;
    int64_t L = 0;                    // leaves of cloud 0 (the single-cloud hierarchy)
    int ncl = 0;
    std::vector<Cloud> hcl;
    int64_t nleaves = 0, Lmax = 0;
    int mbits = 16;                   // Morton bits per axis
    int64_t target_edges = 0;         // sum over clouds of (points - 1)
    const float* xyz = nullptr;       // input points on the device (AoS)
    bool have_points = false;
    bool built = false;
    bool finished = false;
    uint64_t alloc_gen = 0;           // counts the buffer (re)allocations: part of every graph's key
    int rounds = 0;
    uint32_t edges = 0;
    uint32_t zero_or_empty = 0;
    std::vector<emst_round_trace> trace;
    float times[4] = {0, 0, 0, 0};
    int64_t stats[18] = {};
    uint32_t resumed_h = 0;           // queries resumed after the first search pass
    float seed_scale = 3.0f;          // plain seed: multiple of the leaf cell's mean spacing
    hipEvent_t ev[14] = {};             // 8..11: around the component search's dispatches, 12..13:
                                        // around the neighbour lists' dispatch (traced runs)
    float comp_search_ms = 0.f;         // those dispatches' time, summed over the traced run's rounds
    float knn_ms = 0.f;                 // the neighbour lists' dispatch (traced runs)
    // device buffers
    DevBuf<float> xyz_own, resc;
    DevBuf<uint64_t> keys0, keys1, cand_key, corr_key, comp1;
    DevBuf<uint32_t> vals0, vals1, hist, scan_tmp, comp2;
    DevBuf<float4> pts;
    DevBuf<int32_t> inv, cand_pos, label, label2, corr, list, eu, ev_, eround, geo;
    DevBuf<uint32_t> ids;              // input id of each sorted position (the w of pts, dense)
    DevBuf<Node> nodes;
    DevBuf<int32_t> lstart, leaf_of, cloud_lf;     // LeafTable and each cloud's first leaf
    DevBuf<float> ed2;
    DevBuf<double> partial;
    DevBuf<float> model, feat, pred, seed, ls;     // learned seed radius
    DevBuf<Cloud> dcl;
    DevBuf<uint64_t> rkey;             // cone candidates ranked by key, K per point (k_round_min)
    DevBuf<int32_t> rpos;
    DevBuf<uint8_t> cursor;            // per-point scan position in its ranked candidates
    DevBuf<uint64_t> open_mask;        // split cone search (SplitScratch)
    DevBuf<uint32_t> open_cnt, open_tmp;
    DevBuf<Params> params;
    DevBuf<double> cloud_w;
    DevBuf<int32_t> cloud_e, cloud_r;
    bool learned = false;
    // coincident-point groups (DupState)
    DevBuf<int32_t> dup_rep, dup_list, dup_runs, dup_gslot, dup_bpos;
    DevBuf<uint32_t> dup_m2;
    DevBuf<uint64_t> dup_bkey;
    // cycle rule of the 2.5D mode (CycleState)
    DevBuf<int32_t> cy_j0, cy_j, cy_roots;
    DevBuf<uint64_t> cy_kh;
    DevBuf<uint32_t> cy_kl;
    DevBuf<uint8_t> cy_onc, cy_cut;
    DevBuf<unsigned long long> visit;  // arrival state of the bottom-up node passes
    DevBuf<int32_t> list2;             // queries open after pass 0 (per-round component search)
    DevBuf<uint64_t> defer;            // queries handed to the per-query pass (per-round component search)
    DevBuf<int32_t> knn_pos;           // neighbour lists of the component search (NbrList)
    DevBuf<uint64_t> knn_key;
    DevBuf<uint8_t> knn_cur;
    bool knn_built = false;            // the build wrote the lists (not with the learned component seed)
    DevBuf<int32_t> csize;             // component sizes at the roots (learned component seed)
    DevBuf<uint32_t> cbox, nq_keep;    // component boxes at the roots; searching points' count
    DevBuf<float> samples;             // logged (features, target) pairs of the component search
    bool no_grouping = false;         // set when distinct points met at float distance 0
    bool grouped = false;
    bool groups_settled = false;      // no group spans two components any more
    hipGraphExec_t graph = nullptr;   // captured rounds (run_graph)
    std::vector<uintptr_t> graph_key;
    hipGraphExec_t graph1 = nullptr;  // a whole captured run (one_launch)
    std::vector<uintptr_t> graph1_key;
    // one_launch: device clock at the run's start [0], at the end of each
    // round [1..kOneLaunchRounds], after the build [kOneLaunchRounds + 1] and
    // at the end [kOneLaunchRounds + 2]
    DevBuf<unsigned long long> stamps;
    int clock_khz = 0;                 // hipDeviceAttributeWallClockRate
    std::vector<float> round_end_ms;   // the last one-launch run: start to the end of each round
    int64_t dup_points = 0, dup_groups = 0;
    Params* P = nullptr;
    Counters* C = nullptr;
};

namespace {

int fail(emst_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, EMST_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(ctx)                                                                        \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, EMST_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int ensure(emst_ctx* c, DevBuf<T>& b, size_t count)
{
    if (count == 0) count = 1;
    if (b.cap >= count) return EMST_OK;
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc((void**)&b.p, sizeof(T) * count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(c, EMST_E_OOM, "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed");
    }
    b.cap = count;
    ++c->alloc_gen;
    return EMST_OK;
}

template <typename T>
void release(DevBuf<T>& b)
{
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// leaf cells hold a power of two of points in [8, 64] (the cone search maps
// one leaf point per lane, 64 / bucket queries per pass of the wave).  The
// default is the fastest measured at 10^7 points: for the cone search 16 in
// 2D (8 cones, cheap pairs: smaller leaves cut pairs more than they add node
// tests) and 64 in 3D; for the component search 64 in both (2D power-law:
// 21.3 ms per EMST against 22.1 with 16).
int choose_bucket(const emst_options& o, int dim)
{
    int b = o.bucket_size;
    if (b <= 0) b = dim == 2 && o.round_search != 1 ? 16 : 64;
    int r = 8;
    while (r < b && r < 64) r <<= 1;
    return r;
}

int alloc_all(emst_ctx* c)
{
    const int64_t n = c->n;
    const bool comp = c->opt.round_search == 1;
    const int K = comp ? 0 : c->K;      // the component search keeps no cone candidates
    int rc;
#define ENS(buf, cnt) \
    if ((rc = ensure(c, c->buf, (size_t)(cnt))) != EMST_OK) return rc;
    ENS(resc, n * c->dim);
    ENS(keys0, n);
    ENS(keys1, n);
    ENS(vals0, n);
    ENS(vals1, n);
    ENS(hist, radix_hist_words(n));
    ENS(scan_tmp, scan_tmp_words((int64_t)radix_hist_words(n)) + scan_tmp_words(n) + 16);
    ENS(pts, n);
    ENS(inv, n);
    ENS(ids, n);
    ENS(cand_key, n * K);
    ENS(cand_pos, n * K);
    ENS(label, n);
    ENS(label2, n);
    ENS(corr, n);
    ENS(list, n);
    ENS(corr_key, n);
    ENS(comp1, n);
    ENS(comp2, n);
    ENS(eu, n);
    ENS(ev_, n);
    ENS(eround, n);
    ENS(ed2, n);
    ENS(partial, kWeightBlocks);
    ENS(geo, 16);
    ENS(rkey, n * K);
    ENS(rpos, n * K);
    ENS(cursor, n);
    {
        const int64_t P = (n + 63) / 64;
        ENS(open_mask, P);
        ENS(open_cnt, 2 * (P + 1));
        ENS(open_tmp, scan_tmp_words(P + 1) + 16);
    }
    ENS(dup_rep, n);
    ENS(dup_list, n);
    ENS(dup_runs, n);
    ENS(dup_gslot, n);
    if (comp) {
        ENS(list2, n);
        ENS(defer, n);
        ENS(knn_pos, (size_t)knn_len(c->dim) * n);
        ENS(knn_key, (size_t)knn_len(c->dim) * n);
        ENS(knn_cur, n);
    }
    if (c->cdim < c->dim) {
        ENS(cy_j0, n);
        ENS(cy_roots, n);
        ENS(cy_j, 2 * n);
        ENS(cy_kh, 2 * n);
        ENS(cy_kl, 2 * n);
        ENS(cy_onc, n);
        ENS(cy_cut, n);
    }
#undef ENS
    return EMST_OK;
}

int read_counters(emst_ctx* c, Counters* h)
{
    HIPCHK(c, hipMemcpyAsync(h, c->C, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMST_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

static_assert(offsetof(Counters, cycles_cut) + sizeof(uint32_t) - offsetof(Counters, dup_runs) ==
                  8 * sizeof(uint32_t),
              "the dup counters and cycles_cut are cleared as 8 consecutive words");
static_assert(offsetof(Counters, zero_or_empty) == offsetof(Counters, research) + sizeof(uint32_t),
              "research and zero_or_empty are cleared together");

// Cloud positions (hcl[k].n already set) and leaves of at most `bucket`
// points; sizes the buffers.  The leaves themselves, and with them each
// cloud's hierarchy, come from the sorted keys in the build (make_leaves).
int layout_clouds(emst_ctx* c, int bucket)
{
    c->bucket = bucket;
    int64_t pos = 0, target = 0;
    for (Cloud& cl : c->hcl) {
        cl.pos = pos;
        cl.lbase = 0;
        cl.L = 1;
        cl.lfirst = 0;
        cl.nl = 0;
        pos += cl.n;
        target += std::max<int64_t>(cl.n - 1, 0);
    }
    c->target_edges = target;
    // Morton bits per axis: the cloud index takes the top key bits of a batch.
    // Keys of 48 bits (16 per axis in 3D, 24 in 2D, fewer in a batch) sort in
    // 6 radix passes instead of 8; the leaves stay the same cells wherever a
    // cell holds at most a leaf's points, down to 2^-16 of the box (a longer
    // run of equal keys is cut by count, as a run of coincident points is):
    // 10^7 points 0.2 ms faster per EMST in 2D and 3D.  Batches of more than
    // 2^16 clouds use 64 bits.
    int cb = 0;
    if (c->ncl > 1)
        while ((1ll << cb) < c->ncl) ++cb;
    c->mbits = std::min(c->dim == 3 ? 16 : 24, ((cb <= 16 ? 48 : 64) - cb) / c->dim);
    int rc;
    if ((rc = alloc_all(c)) || (rc = ensure(c, c->dcl, (size_t)c->ncl)) ||
        (rc = ensure(c, c->params, (size_t)c->ncl)))
        return rc;
    c->P = c->params.p;
    HIPCHK(c, hipMemcpyAsync(c->dcl.p, c->hcl.data(), sizeof(Cloud) * c->ncl, hipMemcpyHostToDevice, c->stream));
    return EMST_OK;
}

LeafTable leaf_table(const emst_ctx* c) { return LeafTable{c->lstart.p, c->leaf_of.p}; }

// The leaves of the sorted points (LeafTable), then each cloud's hierarchy:
// its leaf count sets its slot count L (a power of two) and its node offset.
// One host synchronisation, for the leaf counts.
int make_leaves(emst_ctx* c)
{
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    int rc;
    if ((rc = ensure(c, c->lstart, (size_t)n + 1)) || (rc = ensure(c, c->leaf_of, (size_t)n)) ||
        (rc = ensure(c, c->cloud_lf, (size_t)c->ncl + 1)))
        return rc;
    // the sort's second key and value buffers are free here
    launch_leaves(c->keys0.p, n, c->dcl.p, c->ncl, c->bucket, c->vals1.p, reinterpret_cast<uint32_t*>(c->keys1.p),
                  c->scan_tmp.p, c->lstart.p, c->leaf_of.p, c->cloud_lf.p, c->P, c->C, false, s);
    LAUNCHCHK(c);
    std::vector<int32_t> lf(c->ncl + 1);
    HIPCHK(c, hipMemcpyAsync(lf.data(), c->cloud_lf.p, sizeof(int32_t) * (c->ncl + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    int64_t lbase = 0, Lmax = 1;
    for (int k = 0; k < c->ncl; ++k) {
        Cloud& cl = c->hcl[k];
        cl.lfirst = lf[k];
        cl.nl = lf[k + 1] - lf[k];
        int64_t L = 1;
        while (L < cl.nl) L <<= 1;
        cl.L = L;
        cl.lbase = lbase;
        lbase += L;
        Lmax = std::max(Lmax, L);
    }
    c->nleaves = lbase;
    c->Lmax = Lmax;
    c->L = c->hcl[0].L;
    if ((rc = ensure(c, c->nodes, 2 * (size_t)lbase))) return rc;
    if (c->opt.round_search == 1 && (rc = ensure(c, c->visit, 2 * (size_t)lbase))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->dcl.p, c->hcl.data(), sizeof(Cloud) * c->ncl, hipMemcpyHostToDevice, s));
    return EMST_OK;
}

CycleState cycle_state(emst_ctx* c)
{
    CycleState y{};
    if (c->cdim >= c->dim) return y;
    const int64_t n = c->n;
    y.j0 = c->cy_j0.p;
    y.roots = c->cy_roots.p;
    y.j[0] = c->cy_j.p;
    y.j[1] = c->cy_j.p + n;
    y.kh[0] = c->cy_kh.p;
    y.kh[1] = c->cy_kh.p + n;
    y.kl[0] = c->cy_kl.p;
    y.kl[1] = c->cy_kl.p + n;
    y.onc = c->cy_onc.p;
    y.cut = c->cy_cut.p;
    y.inv = c->inv.p;
    int steps = 1;
    while (steps < 62 && ((int64_t)1 << steps) < n) ++steps;
    y.steps = steps;
    return y;
}

constexpr int64_t kSampleCap = 1 << 20;         // logged pairs per run
// Default plain seed of the 3D component search, in leaf spacings: its first
// pass looks for one nearest point, not for every cone's.  At 10^7 points 1.0
// beats 3.0 by 10% (clustered) to 15% (uniform) and 1.25-2.0 by 2-6%; in 2D
// 3.0 stays best.
constexpr float kCompSeedScale3D = 1.0f;
constexpr int64_t kSamplesPerRound = 1 << 16;   // at most, per round

// the learned seed of the component search: buffers when a model is set or
// pairs are logged
bool comp_features_on(const emst_ctx* c)
{
    return c->opt.round_search == 1 && (c->learned || c->opt.log_samples == 1);
}

int ensure_comp_seed(emst_ctx* c)
{
    if (!comp_features_on(c)) return EMST_OK;
    const int64_t n = c->n;
    int rc;
    if ((rc = ensure(c, c->csize, (size_t)n)) || (rc = ensure(c, c->cbox, (size_t)n * 6)) ||
        (rc = ensure(c, c->feat, (size_t)n * 8)) || (rc = ensure(c, c->ls, (size_t)n)) ||
        (rc = ensure(c, c->pred, (size_t)n)) || (rc = ensure(c, c->seed, (size_t)n)) ||
        (rc = ensure(c, c->nq_keep, 2)))
        return rc;
    if (c->opt.log_samples == 1 && (rc = ensure(c, c->samples, (size_t)kSampleCap * 9))) return rc;
    return EMST_OK;
}

CompSeed comp_seed(emst_ctx* c)
{
    CompSeed cs{};
    cs.features_on = comp_features_on(c);
    if (!cs.features_on) return cs;
    cs.csize = c->csize.p;
    cs.cbox = c->cbox.p;
    cs.X = c->feat.p;
    cs.LS = c->ls.p;
    cs.Y = c->pred.p;
    cs.R = c->seed.p;
    cs.model = c->learned ? c->model.p : nullptr;
    cs.samples = c->opt.log_samples == 1 ? c->samples.p : nullptr;
    cs.cap = kSampleCap;
    cs.per_round = kSamplesPerRound;
    cs.nq_keep = c->nq_keep.p;
    return cs;
}

// the split-search buffers, also the per-round component search's query list
SplitScratch split_scratch(emst_ctx* c)
{
    SplitScratch sc{};
    sc.open_mask = c->open_mask.p;
    sc.cnt = c->open_cnt.p;
    sc.off = c->open_cnt.p + ((c->n + 63) / 64 + 1);
    sc.tmp = c->open_tmp.p;
    sc.list = c->list.p;
    sc.list2 = c->list2.p;
    sc.defer = c->defer.p;
    sc.after = 0;
    sc.compact_frac = 0.10;
    return sc;
}

DupState dup_state(emst_ctx* c)
{
    DupState d{};
    if (!c->grouped) return d;
    d.rep = c->dup_rep.p;
    d.list = c->dup_list.p;
    d.gslot = c->dup_gslot.p;
    d.m2 = c->dup_m2.p;
    d.bkey = c->dup_bkey.p;
    d.bpos = c->dup_bpos.p;
    d.npoints = c->dup_points;
    d.ngroups = c->dup_groups;
    return d;
}

// bbox, rescale, Morton sort, hierarchy, round-1 cone search
NbrList nbr_list(emst_ctx* c)
{
    if (!c->knn_built) return NbrList{nullptr, nullptr, nullptr};
    return NbrList{c->knn_pos.p, c->knn_key.p, c->knn_cur.p};
}

int build(emst_ctx* c)
{
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    launch_bbox(c->xyz, c->dcl.p, c->ncl, n, c->dim, c->P, s);
    launch_params(c->P, c->dcl.p, c->ncl, c->dim, c->mbits, s);
    LAUNCHCHK(c);
    {
        std::vector<Params> hp(c->ncl);
        HIPCHK(c, hipMemcpyAsync(hp.data(), c->P, sizeof(Params) * c->ncl, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (int k = 0; k < c->ncl; ++k)
            if (hp[k].nonfinite)
                return fail(c, EMST_E_NONFINITE,
                            c->ncl == 1 ? std::string("input holds a NaN or Inf coordinate")
                                        : "cloud " + std::to_string(k) + " holds a NaN or Inf coordinate");
    }
    launch_rescale_morton(c->xyz, c->dcl.p, c->ncl, n, c->dim, c->mbits, c->P, c->resc.p, c->keys0.p, s);
    int key_bits = c->mbits * c->dim;
    if (c->ncl > 1) {
        int cb = 0;
        while ((1ll << cb) < c->ncl) ++cb;
        key_bits += cb;
    }
    radix_sort(c->keys0.p, c->vals0.p, c->keys1.p, c->vals1.p, n, key_bits, true, c->hist.p, c->scan_tmp.p, s);
    launch_gather_sorted(c->resc.p, c->vals0.p, n, c->dim, c->pts.p, c->inv.p, c->ids.p, s);
    int rcl = make_leaves(c);
    if (rcl) return rcl;
    // coincident-point groups: dup_runs .. cycles_cut start at zero
    HIPCHK(c, hipMemsetAsync(&c->C->dup_runs, 0, 8 * sizeof(uint32_t), s));
    c->grouped = false;
    c->groups_settled = false;
    c->knn_built = false;
    c->dup_points = c->dup_groups = 0;
    const bool comp = c->opt.round_search == 1;
    if (!c->no_grouping && !comp) {
        launch_dup_groups(c->keys0.p, c->pts.p, n, c->dup_rep.p, c->dup_runs.p, c->dup_list.p,
                          c->dup_gslot.p, c->C, s);
        LAUNCHCHK(c);
        Counters h{};
        int rc0 = read_counters(c, &h);
        if (rc0) return rc0;
        if (h.grouped) {
            const int K2 = c->K / 2;
            if ((rc0 = ensure(c, c->dup_m2, h.dup_groups)) ||
                (rc0 = ensure(c, c->dup_bkey, (size_t)K2 * h.dup_groups)) ||
                (rc0 = ensure(c, c->dup_bpos, (size_t)K2 * h.dup_groups)))
                return rc0;
            c->grouped = true;
            c->dup_points = h.dup_points;
            c->dup_groups = h.dup_groups;
        }
    }
    const DupState dup = dup_state(c);
    launch_build_tree(c->pts.p, c->dcl.p, c->ncl, c->nleaves, c->Lmax, leaf_table(c), c->nodes.p, dup.rep, s);
    launch_round_init(n, c->label.p, c->label2.p, c->eu.p, c->corr.p, c->corr_key.p, c->comp1.p,
                      c->comp2.p, c->C, (uint32_t)c->target_edges, (uint32_t)c->opt.max_rounds, s);
    LAUNCHCHK(c);
    if (comp) {
        // the per-round component search has no round-1 search: every round
        // searches (launch_comp_round)
        HIPCHK(c, hipMemsetAsync(c->visit.p, 0, sizeof(unsigned long long) * 2 * c->nleaves, s));
        launch_node_min_ids(c->pts.p, c->dcl.p, c->ncl, c->nleaves, leaf_table(c), c->nodes.p, c->visit.p, s);
        // round 1's targets and the neighbour lists; a run that logs the
        // learned seed's training pairs searches in round 1 as well (its log
        // starts there); the seed itself applies from round 3 on, past the
        // rounds the lists answer
        c->knn_built = c->opt.log_samples != 1;
        const bool tr = c->opt.trace != 0;
        if (tr) HIPCHK(c, hipEventRecord(c->ev[12], s));
        if (c->knn_built)
            launch_knn_first(c->dim, n, c->pts.p, c->nodes.p, c->dcl.p, c->ncl, c->bucket, leaf_table(c), c->inv.p,
                             c->corr.p, c->corr_key.p, nbr_list(c), c->C, tr, s);
        if (tr) HIPCHK(c, hipEventRecord(c->ev[13], s));
        int rc1 = ensure_comp_seed(c);
        if (rc1) return rc1;
        if (comp_features_on(c)) launch_comp_seed_init(n, c->pts.p, c->csize.p, c->cbox.p, s);
        LAUNCHCHK(c);
        HIPCHK(c, hipEventRecord(c->ev[1], s));
        HIPCHK(c, hipEventRecord(c->ev[2], s));
        HIPCHK(c, hipEventRecord(c->ev[3], s));
        Counters h{};
        int rc = read_counters(c, &h);
        if (rc) return rc;
        c->zero_or_empty = 0;
        c->comp_search_ms = 0.f;
        for (auto& x : c->stats) x = 0;
        c->resumed_h = 0;
        c->times[0] = elapsed(c->ev[0], c->ev[1]);
        c->times[1] = 0.f;
        c->knn_ms = tr ? elapsed(c->ev[12], c->ev[13]) : 0.f;
        c->built = true;
        c->rounds = 0;
        c->edges = 0;
        c->finished = c->target_edges == 0;
        c->trace.clear();
        return EMST_OK;
    }
    const float* seed_r = nullptr;
    if (c->learned && c->ncl == 1 && c->cdim == c->dim) {
        int rc2;
        if ((rc2 = ensure(c, c->feat, (size_t)n * 8)) || (rc2 = ensure(c, c->pred, (size_t)n)) ||
            (rc2 = ensure(c, c->seed, (size_t)n)) || (rc2 = ensure(c, c->ls, (size_t)n)))
            return rc2;
        launch_radius_features(c->dim, c->pts.p, c->nodes.p, c->L, leaf_table(c), n, c->feat.p, c->ls.p, s);
        launch_radius_seed(n, c->model.p, c->feat.p, c->ls.p, c->pred.p, c->seed.p, s);
        seed_r = c->seed.p;
    }
    LAUNCHCHK(c);
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    // 3D cones over 2^21 points or more: pass 0 over all queries, the open
    // ones then resumed from a compacted list when at most 10% remain, else
    // in place.  Measured at 10^7 points: 8% open on the clustered cloud
    // gains 5%, 7% on a uniform one 6%.  At 10^6 points (12% open) the second
    // launch costs more than it saves (-5%), and in 2D (8 cheap cones,
    // 16-point leaves) the single launch is faster at every open share
    // measured (10-26%).
    SplitScratch split{};
    split.open_mask = c->open_mask.p;
    split.cnt = c->open_cnt.p;
    split.off = c->open_cnt.p + ((n + 63) / 64 + 1);
    split.tmp = c->open_tmp.p;
    split.list = c->list.p;     // free until the rounds' research queue
    split.after = 0;
    const int mode = c->opt.search_split;
    split.compact_frac = mode == 2 ? -1.0 : (mode == 3 ? 1.0 : 0.10);
    const bool use_split = mode == 2 || mode == 3 || (mode == 0 && c->cdim == 3 && n >= ((int64_t)1 << 21));
    launch_cone_search(c->dim, c->cdim, c->pts.p, c->nodes.p, c->dcl.p, c->ncl, c->bucket, leaf_table(c), n, nullptr, n,
                       c->label.p, false, c->inv.p, seed_r, c->seed_scale, c->cand_key.p, c->cand_pos.p, nullptr, c->corr.p,
                       c->corr_key.p, c->C, true, dup, use_split ? &split : nullptr, s);
    LAUNCHCHK(c);
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    // read with the counters below (into the context, which outlives any
    // early return while the copy is queued)
    c->resumed_h = 0;
    if (use_split && n > 0)
        HIPCHK(c, hipMemcpyAsync(&c->resumed_h, split.off + (n + 63) / 64, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 s));
    launch_dup_fill(c->cdim, n, c->label.p, c->pts.p, c->inv.p, c->cand_key.p, c->cand_pos.p, nullptr, c->corr.p,
                    c->corr_key.p, c->C, dup, s);
    launch_rank_cones(c->cdim, n, c->cand_key.p, c->cand_pos.p, nullptr, n, c->rkey.p, c->rpos.p, c->cursor.p,
                      nullptr, s);
    LAUNCHCHK(c);
    // while grouping is on only group points can hold a zero-distance or no
    // candidate, and k_dup_fill refreshes those; the queue serves the
    // per-point form
    if (!c->grouped) launch_flag_research(c->K, n, c->cand_key.p, c->list.p, c->C, s);
    LAUNCHCHK(c);
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    Counters h{};
    int rc = read_counters(c, &h);
    if (rc) return rc;
    if (c->grouped && h.underflow) {
        // distinct points at float distance 0: the groups would merge
        // candidates that the per-point rule keeps apart; search ungrouped
        c->no_grouping = true;
        return build(c);
    }
    c->zero_or_empty = h.zero_or_empty;
    c->stats[0] = (int64_t)h.packets;
    c->stats[1] = (int64_t)h.nodes;
    c->stats[2] = (int64_t)h.leaves;
    c->stats[3] = (int64_t)h.leaf_points;
    c->stats[4] = (int64_t)h.pairs;
    c->stats[5] = (int64_t)h.dup_points;
    c->stats[6] = c->grouped ? 1 : 0;
    c->stats[7] = 0;
    c->stats[8] = c->resumed_h;
    if (c->grouped) c->groups_settled = h.dup_split == 0u;
    c->times[0] = elapsed(c->ev[0], c->ev[1]) + elapsed(c->ev[2], c->ev[3]);
    c->times[1] = elapsed(c->ev[1], c->ev[2]);
    c->built = true;
    c->rounds = 0;
    c->edges = 0;
    c->finished = c->target_edges == 0;
    c->trace.clear();
    return EMST_OK;
}

int weight_of(emst_ctx* c, int64_t m, double* w)
{
    *w = 0;
    if (m <= 0) return EMST_OK;
    launch_weight(c->xyz, c->dim, c->eu.p, c->ev_.p, m, c->partial.p, kWeightBlocks, c->stream);
    LAUNCHCHK(c);
    std::vector<double> h(kWeightBlocks);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial.p, sizeof(double) * kWeightBlocks, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s = 0;
    for (double x : h) s += x;
    *w = s;
    return EMST_OK;
}

// rounds captured in a one-launch run (options.one_launch); a run that needs
// more finishes with the replayed round graph.  Rounds past the end cost
// their launches (about 0.15 ms each at 10^7 points), and 10^7-point clouds
// finish in 10-12 rounds.
constexpr int kOneLaunchRounds = 16;

// The kernels of one Boruvka round (SolutionEMST::activate).  Round control
// lives on the device (k_round_begin / k_round_end), so a round issued after
// the run has finished does nothing.  trace: events after each phase.
void issue_round(emst_ctx* c, bool tr, unsigned long long* stamps = nullptr)
{
    hipStream_t s = c->stream;
    launch_round_begin(c->C, s);
    if (tr) (void)hipEventRecord(c->ev[3], s);
    // group points: zero-distance '+' cones while the group spans several
    // components, the representative's own cones once it is one (the
    // repeated search of NodeEMST.h:1051-1077, without searching again); the
    // kernels skip themselves once the groups have settled
    if (c->grouped) {
        const DupState d = dup_state(c);
        launch_dup_fill(c->cdim, c->n, c->label.p, c->pts.p, c->inv.p, c->cand_key.p, c->cand_pos.p, c->corr.p,
                        nullptr, nullptr, c->C, d, s);
        launch_rank_cones(c->cdim, c->n, c->cand_key.p, c->cand_pos.p, d.list, d.npoints, c->rkey.p, c->rpos.p,
                          c->cursor.p, c->C, s);
    }
    const CycleState cy = cycle_state(c);
    for (int ph = 0; ph < 4; ++ph) {
        if (ph == 0 && c->opt.round_search == 1) {
            launch_comp_round(c->dim, c->n, c->pts.p, c->ids.p, c->nodes.p, c->visit.p, c->dcl.p, c->ncl,
                              c->nleaves, c->bucket, leaf_table(c), c->label.p, c->corr.p, c->corr_key.p, c->comp1.p,
                              c->inv.p, c->seed_scale, split_scratch(c), comp_seed(c), nbr_list(c), c->eu.p,
                              c->eround.p, c->C, s, tr ? c->ev + 8 : nullptr);
            if (tr) (void)hipEventRecord(c->ev[4], s);
            continue;
        }
        launch_round_phase(ph, c->K, c->n, c->ids.p, c->rkey.p, c->rpos.p, c->cursor.p,
                           c->label.p, c->label2.p,
                           c->corr.p, c->corr_key.p, c->comp1.p, c->comp2.p, c->C, c->eu.p,
                           c->ev_.p, c->ed2.p, c->eround.p, cy, s);
        if (tr) (void)hipEventRecord(c->ev[4 + ph], s);
    }
    launch_round_end(c->C, c->grouped, s, stamps, kOneLaunchRounds);
}

// A flatten that stopped at its step cap paused the rounds (done == 2):
// launch it until the labels are flat, then end the round again.
int resume_flatten(emst_ctx* c, Counters* h)
{
    hipStream_t s = c->stream;
    int rc;
    for (int again = 0; h->unflat; ++again) {
        if (again == 64) return fail(c, EMST_E_HIP, "flatten did not converge (cycle in the hook forest)");
        HIPCHK(c, hipMemsetAsync(&c->C->unflat, 0, sizeof(uint32_t), s));
        launch_round_phase(3, c->K, c->n, c->ids.p, c->rkey.p, c->rpos.p, c->cursor.p,
                           c->label.p, c->label2.p,
                           c->corr.p, c->corr_key.p, c->comp1.p, c->comp2.p, c->C, c->eu.p,
                           c->ev_.p, c->ed2.p, c->eround.p, CycleState{}, s);
        LAUNCHCHK(c);
        if ((rc = read_counters(c, h))) return rc;
    }
    HIPCHK(c, hipMemsetAsync(&c->C->done, 0, sizeof(uint32_t), s));
    launch_round_end(c->C, c->grouped, s);
    LAUNCHCHK(c);
    return read_counters(c, h);
}

void absorb_counters(emst_ctx* c, const Counters& h)
{
    c->edges = h.edges;
    c->rounds = (int)h.round;
    c->stats[7] = (int64_t)h.cycles_cut;
    c->stats[9] = (int64_t)h.comp_queries;
    for (int k = 0; k < kWorkKinds; ++k) {
        unsigned long long t = 0;
        for (int j = 0; j < kWorkSlots; ++j) t += h.work[k][j];
        c->stats[10 + k] = (int64_t)t;
    }
    if (h.done == 1u) c->finished = true;
}

// One round from the host (emst_step, traced runs, and runs that repeat cone
// searches).  Termination is tested first.
int round_once(emst_ctx* c, int32_t* finished)
{
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    if (c->finished || (int64_t)c->edges >= c->target_edges || c->rounds >= c->opt.max_rounds) {
        c->finished = true;
        *finished = 1;
        return EMST_OK;
    }
    *finished = 0;
    const bool tr = c->opt.trace != 0;
    int64_t searched = c->rounds == 0 ? n : 0;
    if (c->rounds > 0 && c->grouped) {
        if (!c->groups_settled) searched = c->dup_points;
    } else if (c->rounds > 0 && c->zero_or_empty > 0) {
        // research and zero_or_empty restart for this round
        HIPCHK(c, hipMemsetAsync(&c->C->research, 0, 2 * sizeof(uint32_t), s));
        launch_flag_research(c->K, n, c->cand_key.p, c->list.p, c->C, s);
        LAUNCHCHK(c);
        Counters h{};
        int rc = read_counters(c, &h);
        if (rc) return rc;
        if (h.research > 0) {
            launch_cone_search(c->dim, c->cdim, c->pts.p, c->nodes.p, c->dcl.p, c->ncl, c->bucket, leaf_table(c), n, c->list.p,
                               h.research, c->label.p, true, c->inv.p, nullptr, c->seed_scale, c->cand_key.p,
                               c->cand_pos.p, c->corr.p, nullptr, nullptr, c->C, false, DupState{}, nullptr, s);
            launch_rank_cones(c->cdim, n, c->cand_key.p, c->cand_pos.p, c->list.p, h.research, c->rkey.p, c->rpos.p,
                              c->cursor.p, nullptr, s);
        }
        LAUNCHCHK(c);
        searched = h.zero_or_empty;
    }
    const int64_t queries_before = c->stats[9];
    issue_round(c, tr);
    LAUNCHCHK(c);
    Counters h{};
    int rc = read_counters(c, &h);
    if (rc) return rc;
    if (tr && c->opt.round_search == 1 && n > 0)
        c->comp_search_ms += elapsed(c->ev[8], c->ev[9]) + elapsed(c->ev[10], c->ev[11]);
    if (h.done == 2u && (rc = resume_flatten(c, &h))) return rc;
    absorb_counters(c, h);
    // the component search: queries of this round's two dispatches
    if (c->opt.round_search == 1) searched = c->stats[9] - queries_before;
    c->groups_settled = h.dup_settled != 0u;
    if (tr) {
        emst_round_trace row{};
        row.iteration = c->rounds;
        row.size = (int32_t)c->edges;
        double w = 0;
        rc = weight_of(c, n, &w);
        if (rc) return rc;
        row.length = w;
        row.time_find_next_closest = elapsed(c->ev[3], c->ev[4]);
        row.time_find_min_pair = elapsed(c->ev[4], c->ev[5]);
        row.time_connect_union = elapsed(c->ev[5], c->ev[6]);
        row.time_flatten = elapsed(c->ev[6], c->ev[7]);
        row.time_termination = 0;
        row.searched = searched;
        c->trace.push_back(row);
    }
    return EMST_OK;
}

// Untraced runs without repeated cone searches: kGraphRounds rounds captured
// once in a graph and replayed, one host synchronisation per replay; rounds
// past the end do nothing.  The graph is kept while the buffers and sizes it
// captured stay the same.
constexpr int kGraphRounds = 4;

std::vector<uintptr_t> graph_key(emst_ctx* c)
{
    const DupState d = dup_state(c);
    const CycleState y = cycle_state(c);
    return {(uintptr_t)c->n, (uintptr_t)c->K, (uintptr_t)c->dim, (uintptr_t)c->cdim, (uintptr_t)y.j0, (uintptr_t)y.roots,
            (uintptr_t)c->opt.round_search, (uintptr_t)comp_features_on(c), (uintptr_t)c->learned,
            (uintptr_t)c->model.p, (uintptr_t)c->csize.p, (uintptr_t)c->cbox.p, (uintptr_t)c->feat.p, (uintptr_t)c->ls.p,
            (uintptr_t)c->pred.p, (uintptr_t)c->seed.p, (uintptr_t)c->samples.p, (uintptr_t)c->nq_keep.p,
            (uintptr_t)c->opt.log_samples, (uintptr_t)c->visit.p, (uintptr_t)c->list2.p, (uintptr_t)c->defer.p, (uintptr_t)c->knn_pos.p,
            (uintptr_t)c->knn_key.p, (uintptr_t)c->knn_cur.p, (uintptr_t)c->knn_built, (uintptr_t)c->open_mask.p, (uintptr_t)c->open_cnt.p,
            (uintptr_t)c->open_tmp.p, (uintptr_t)c->list.p, (uintptr_t)c->nodes.p, (uintptr_t)c->dcl.p,
            (uintptr_t)c->nleaves, (uintptr_t)c->lstart.p, (uintptr_t)c->leaf_of.p, (uintptr_t)c->bucket,
            (uintptr_t)__builtin_bit_cast(uint32_t, c->seed_scale),
            (uintptr_t)y.j[0], (uintptr_t)y.kh[0], (uintptr_t)y.kl[0], (uintptr_t)y.onc, (uintptr_t)y.cut,
            (uintptr_t)c->grouped, (uintptr_t)d.npoints,
            (uintptr_t)d.ngroups, (uintptr_t)c->C, (uintptr_t)c->pts.p, (uintptr_t)c->cand_key.p,
            (uintptr_t)c->cand_pos.p, (uintptr_t)c->label.p, (uintptr_t)c->label2.p, (uintptr_t)c->corr.p,
            (uintptr_t)c->corr_key.p, (uintptr_t)c->comp1.p, (uintptr_t)c->comp2.p,
            (uintptr_t)c->eu.p, (uintptr_t)c->ev_.p, (uintptr_t)c->ed2.p, (uintptr_t)c->eround.p,
            (uintptr_t)c->inv.p, (uintptr_t)c->rkey.p, (uintptr_t)c->rpos.p, (uintptr_t)c->cursor.p, (uintptr_t)d.rep, (uintptr_t)d.list, (uintptr_t)d.gslot, (uintptr_t)d.m2,
            (uintptr_t)d.bkey, (uintptr_t)d.bpos, (uintptr_t)c->ids.p, (uintptr_t)c->alloc_gen};
}

int run_graph(emst_ctx* c)
{
    hipStream_t s = c->stream;
    const std::vector<uintptr_t> key = graph_key(c);
    if (!c->graph || key != c->graph_key) {
        if (c->graph) {
            (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
        }
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < kGraphRounds; ++r) issue_round(c, false);
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->graph = nullptr;
            return fail(c, EMST_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        c->graph_key = key;
    }
    while (!c->finished) {
        HIPCHK(c, hipGraphLaunch(c->graph, s));
        Counters h{};
        int rc = read_counters(c, &h);
        if (rc) return rc;
        if (h.done == 2u && (rc = resume_flatten(c, &h))) return rc;
        absorb_counters(c, h);
        c->groups_settled = h.dup_settled != 0u;
    }
    return EMST_OK;
}

// ---- one launch: the whole run as one captured graph (options.one_launch)

bool one_launch_on(const emst_ctx* c)
{
    return c->opt.one_launch == 1 && c->opt.round_search == 1 && !c->opt.trace && !comp_features_on(c) &&
           c->n > 0;
}

// The build of the component search with no host decision in it: each
// cloud's slot capacity is fixed from its size, a power of two of at least
// one slot per bucket/4 points (maximal prefix cells of at most 64 points
// hold 25-50 points on average on uniform, clustered and power-law clouds of
// 10^2 to 3.10^5 points; a cloud with more leaves than that is flagged on the
// device and the run redone eagerly).  The device then lays the hierarchies
// out inside that capacity exactly as the eager build does (k_cloud_leaves),
// and the launches over leaf slots cover the capacity, their threads past a
// cloud's slots idle.  The checks the eager build makes on the host (NaN or
// Inf input) leave the rounds idle through C->invalid.
int one_launch_layout(emst_ctx* c)
{
    int64_t lbase = 0, Lmax = 1;
    const int64_t per_slot = std::max(1, c->bucket / 4);
    for (Cloud& cl : c->hcl) {
        const int64_t want = std::max<int64_t>(1, (cl.n + per_slot - 1) / per_slot);
        int64_t L = 1;
        while (L < want) L <<= 1;
        cl.L = L;
        cl.lbase = lbase;
        cl.lfirst = 0;
        cl.nl = 0;
        lbase += L;
        Lmax = std::max(Lmax, L);
    }
    c->nleaves = lbase;
    c->Lmax = Lmax;
    c->L = c->hcl[0].L;
    int rc;
    if ((rc = ensure(c, c->nodes, 2 * (size_t)lbase)) || (rc = ensure(c, c->visit, 2 * (size_t)lbase)) ||
        (rc = ensure(c, c->lstart, (size_t)c->n + 1)) || (rc = ensure(c, c->leaf_of, (size_t)c->n)) ||
        (rc = ensure(c, c->cloud_lf, (size_t)c->ncl + 1)) || (rc = ensure(c, c->stamps, kOneLaunchRounds + 3)))
        return rc;
    if (c->clock_khz <= 0) {
        HIPCHK(c, hipDeviceGetAttribute(&c->clock_khz, hipDeviceAttributeWallClockRate, c->opt.device));
        if (c->clock_khz <= 0) return fail(c, EMST_E_HIP, "no device clock rate");
    }
    HIPCHK(c, hipMemcpyAsync(c->dcl.p, c->hcl.data(), sizeof(Cloud) * c->ncl, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMST_OK;
}

void issue_one_launch_build(emst_ctx* c)
{
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    launch_bbox(c->xyz, c->dcl.p, c->ncl, n, c->dim, c->P, s);
    launch_params(c->P, c->dcl.p, c->ncl, c->dim, c->mbits, s);
    launch_rescale_morton(c->xyz, c->dcl.p, c->ncl, n, c->dim, c->mbits, c->P, c->resc.p, c->keys0.p, s);
    int key_bits = c->mbits * c->dim;
    if (c->ncl > 1) {
        int cb = 0;
        while ((1ll << cb) < c->ncl) ++cb;
        key_bits += cb;
    }
    radix_sort(c->keys0.p, c->vals0.p, c->keys1.p, c->vals1.p, n, key_bits, true, c->hist.p, c->scan_tmp.p, s);
    launch_gather_sorted(c->resc.p, c->vals0.p, n, c->dim, c->pts.p, c->inv.p, c->ids.p, s);
    launch_leaves(c->keys0.p, n, c->dcl.p, c->ncl, c->bucket, c->vals1.p, reinterpret_cast<uint32_t*>(c->keys1.p),
                  c->scan_tmp.p, c->lstart.p, c->leaf_of.p, c->cloud_lf.p, c->P, c->C, true, s);
    // dup_runs .. cycles_cut start at zero (a kernel: no memset nodes in graphs)
    launch_zero_u32(&c->C->dup_runs, 8, s);
    launch_build_tree(c->pts.p, c->dcl.p, c->ncl, c->nleaves, c->Lmax, leaf_table(c), c->nodes.p, nullptr, s);
    launch_round_init(n, c->label.p, c->label2.p, c->eu.p, c->corr.p, c->corr_key.p, c->comp1.p,
                      c->comp2.p, c->C, (uint32_t)c->target_edges, (uint32_t)c->opt.max_rounds, s);
    launch_zero_u32(reinterpret_cast<uint32_t*>(c->visit.p), 4 * c->nleaves, s);
    launch_node_min_ids(c->pts.p, c->dcl.p, c->ncl, c->nleaves, leaf_table(c), c->nodes.p, c->visit.p, s);
    launch_knn_first(c->dim, n, c->pts.p, c->nodes.p, c->dcl.p, c->ncl, c->bucket, leaf_table(c), c->inv.p,
                     c->corr.p, c->corr_key.p, NbrList{c->knn_pos.p, c->knn_key.p, c->knn_cur.p}, c->C, false, s);
}

int run_rounds(emst_ctx* c);

int run_one_launch(emst_ctx* c)
{
    hipStream_t s = c->stream;
    int rc = one_launch_layout(c);
    if (rc) return rc;
    c->grouped = false;
    c->groups_settled = false;
    c->no_grouping = false;
    c->knn_built = true;
    c->dup_points = c->dup_groups = 0;
    std::vector<uintptr_t> key = graph_key(c);
    key.push_back((uintptr_t)c->xyz);
    key.push_back((uintptr_t)c->ncl);
    key.push_back((uintptr_t)c->resc.p);
    key.push_back((uintptr_t)c->keys0.p);
    key.push_back((uintptr_t)c->target_edges);
    key.push_back((uintptr_t)c->opt.max_rounds);
    key.push_back((uintptr_t)c->mbits);
    if (!c->graph1 || key != c->graph1_key) {
        if (c->graph1) {
            (void)hipGraphExecDestroy(c->graph1);
            c->graph1 = nullptr;
        }
        unsigned long long* st = c->stamps.p;
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch_stamp(st, s);
        issue_one_launch_build(c);
        launch_stamp(st + kOneLaunchRounds + 1, s);
        for (int r = 0; r < kOneLaunchRounds; ++r) issue_round(c, false, st);
        launch_stamp(st + kOneLaunchRounds + 2, s);
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&c->graph1, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->graph1 = nullptr;
            return fail(c, EMST_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        c->graph1_key = key;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    HIPCHK(c, hipGraphLaunch(c->graph1, s));
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    Counters h{};
    if ((rc = read_counters(c, &h))) return rc;
    if (h.invalid) {
        std::vector<Params> hp(c->ncl);
        HIPCHK(c, hipMemcpyAsync(hp.data(), c->P, sizeof(Params) * c->ncl, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (int k = 0; k < c->ncl; ++k)
            if (hp[k].nonfinite) {
                c->built = false;
                return fail(c, EMST_E_NONFINITE,
                            c->ncl == 1 ? std::string("input holds a NaN or Inf coordinate")
                                        : "cloud " + std::to_string(k) + " holds a NaN or Inf coordinate");
            }
        // more leaves than slots: the eager build lays the hierarchy out exactly
        c->built = false;
        if ((rc = build(c))) return rc;
        return run_rounds(c);
    }
    c->zero_or_empty = 0;
    c->comp_search_ms = 0.f;
    c->knn_ms = 0.f;
    for (auto& x : c->stats) x = 0;
    c->resumed_h = 0;
    c->built = true;
    c->rounds = 0;
    c->edges = 0;
    c->finished = c->target_edges == 0;
    c->trace.clear();
    if (h.done == 2u && (rc = resume_flatten(c, &h))) return rc;
    absorb_counters(c, h);
    unsigned long long st[kOneLaunchRounds + 3];
    HIPCHK(c, hipMemcpyAsync(st, c->stamps.p, sizeof st, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const double per_ms = (double)c->clock_khz;
    const float whole = elapsed(c->ev[0], c->ev[3]);
    c->times[0] = (float)((st[kOneLaunchRounds + 1] - st[0]) / per_ms);
    c->times[1] = 0.f;
    c->times[2] = std::max(0.f, whole - c->times[0]);
    c->round_end_ms.clear();
    for (int r = 1; r <= std::min(c->rounds, kOneLaunchRounds); ++r)
        c->round_end_ms.push_back((float)((st[r] - st[0]) / per_ms));
    if (!c->finished) {
        HIPCHK(c, hipEventRecord(c->ev[2], s));
        if ((rc = run_graph(c))) return rc;
        HIPCHK(c, hipEventRecord(c->ev[3], s));
        HIPCHK(c, hipEventSynchronize(c->ev[3]));
        c->times[2] += elapsed(c->ev[2], c->ev[3]);
    }
    c->times[3] = c->times[0] + c->times[1] + c->times[2];
    return EMST_OK;
}

int need_points(emst_ctx* c)
{
    if (!c) return EMST_E_ARG;
    if (!c->have_points) return fail(c, EMST_E_STATE, "no points set");
    return EMST_OK;
}

int ensure_built(emst_ctx* c)
{
    int rc = need_points(c);
    if (rc) return rc;
    if (!c->built) return build(c);
    return EMST_OK;
}

// The rounds of a built context, to the end; times[2] = their device time.
int run_rounds(emst_ctx* c)
{
    int rc;
    c->round_end_ms.clear();
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    if (!c->opt.trace && (c->grouped || c->zero_or_empty == 0)) {
        rc = run_graph(c);
        if (rc) return rc;
    } else {
        int32_t fin = 0;
        while (!fin) {
            rc = round_once(c, &fin);
            if (rc) return rc;
        }
    }
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[3]));
    c->times[2] = elapsed(c->ev[2], c->ev[3]);
    c->times[3] = c->times[0] + c->times[1] + c->times[2];
    return EMST_OK;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void emst_default_options(emst_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->dim = 3;
    o->level_radius = 1;
    o->max_rounds = 100;
    o->bucket_size = 0;
    o->trace = 0;
    o->device = 0;
    o->radius_mode = 0;
}

emst_ctx* emst_create(const emst_options* opt)
{
    emst_options o;
    if (opt) o = *opt;
    else emst_default_options(&o);
    if (o.dim != 2 && o.dim != 3) { g_create_error = "dim must be 2 or 3"; return nullptr; }
    if (o.round_search != 0 && o.round_search != 1) {
        g_create_error = "round_search must be 0 or 1";
        return nullptr;
    }
    if (o.round_search == 1 && o.cone_dim == 2 && o.dim == 3) {
        g_create_error = "the 2.5D mode needs the cone search (round_search 0)";
        return nullptr;
    }
    if (o.log_samples != 0 && o.log_samples != 1) {
        g_create_error = "log_samples must be 0 or 1";
        return nullptr;
    }
    if (o.search_split < 0 || o.search_split > 3) {
        g_create_error = "search_split must be 0..3";
        return nullptr;
    }
    if (o.cone_dim != 0 && o.cone_dim != 2 && o.cone_dim != 3) {
        g_create_error = "cone_dim must be 0 (the point dimension), 2 or 3";
        return nullptr;
    }
    if (o.cone_dim == 2 && o.dim == 3 && o.radius_mode != 0) {
        g_create_error = "the learned seed radius is not defined for the 2.5D mode";
        return nullptr;
    }
    if (o.max_rounds <= 0) o.max_rounds = 100;
    if (o.level_radius <= 0) o.level_radius = 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        g_create_error = "no HIP device";
        return nullptr;
    }
    if (o.device < 0 || o.device >= ndev) { g_create_error = "device ordinal out of range"; return nullptr; }
    if (hipSetDevice(o.device) != hipSuccess) { g_create_error = "hipSetDevice failed"; return nullptr; }
    emst_ctx* c = new emst_ctx();
    c->opt = o;
    if (o.seed_scale > 0.f)
        c->seed_scale = o.seed_scale;
    else if (o.round_search == 1 && o.dim == 3)
        c->seed_scale = kCompSeedScale3D;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        delete c;
        return nullptr;
    }
    for (auto& e : c->ev) hipEventCreate(&e);
    if (hipMalloc((void**)&c->C, sizeof(Counters)) != hipSuccess) {
        g_create_error = "hipMalloc of context state failed";
        emst_destroy(c);
        return nullptr;
    }
    return c;
}

int emst_destroy(emst_ctx* c)
{
    if (!c) return EMST_E_ARG;
    hipSetDevice(c->opt.device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->graph) hipGraphExecDestroy(c->graph);
    if (c->graph1) hipGraphExecDestroy(c->graph1);
    if (c->C) hipFree(c->C);
    for (auto& e : c->ev) if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return EMST_OK;
}

const char* emst_last_error(const emst_ctx* c)
{
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int emst_set_clouds(emst_ctx* c, int32_t ncl, const int64_t* npts, const float* xyz, int32_t dim,
                    int32_t on_device)
{
    if (!c) return EMST_E_ARG;
    if (ncl <= 0 || !npts) return fail(c, EMST_E_ARG, "need at least one cloud");
    if (dim != 2 && dim != 3) return fail(c, EMST_E_ARG, "dim must be 2 or 3");
    int64_t n = 0;
    for (int32_t k = 0; k < ncl; ++k) {
        if (npts[k] < 0) return fail(c, EMST_E_ARG, "negative cloud size");
        n += npts[k];
    }
    if (n > 0 && !xyz) return fail(c, EMST_E_ARG, "NULL points");
    if (n >= ((int64_t)1 << 31) - 1) return fail(c, EMST_E_LIMIT, "total points must be below 2^31 - 1");
    const int cdim = c->opt.cone_dim == 0 ? dim : c->opt.cone_dim;
    if (cdim > dim) return fail(c, EMST_E_ARG, "cone_dim must not exceed the point dimension");
    HIPCHK(c, hipSetDevice(c->opt.device));
    c->n = n;
    c->dim = dim;
    c->cdim = cdim;
    c->K = 2 * cdim * (1 << (cdim - 1));
    c->ncl = ncl;
    c->hcl.assign(ncl, Cloud{});
    for (int32_t k = 0; k < ncl; ++k) c->hcl[k].n = npts[k];
    int rc = layout_clouds(c, choose_bucket(c->opt, dim));
    if (rc) return rc;
    if (on_device) {
        c->xyz = xyz;
    } else {
        rc = ensure(c, c->xyz_own, (size_t)(n * dim));
        if (rc) return rc;
        if (n > 0)
            HIPCHK(c, hipMemcpyAsync(c->xyz_own.p, xyz, sizeof(float) * n * dim, hipMemcpyHostToDevice, c->stream));
        c->xyz = c->xyz_own.p;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_points = true;
    c->built = false;
    c->no_grouping = false;
    c->finished = false;
    c->rounds = 0;
    c->edges = 0;
    c->trace.clear();
    return EMST_OK;
}

int emst_set_points(emst_ctx* c, const float* xyz, int64_t n, int32_t dim, int32_t on_device)
{
    if (!c) return EMST_E_ARG;
    if (n < 0 || (n > 0 && !xyz)) return fail(c, EMST_E_ARG, "n < 0 or NULL points");
    return emst_set_clouds(c, 1, &n, xyz, dim, on_device);
}

int emst_autotune(emst_ctx* c, int32_t* best_bucket, float* best_scale, float* best_ms)
{
    int rc = need_points(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->opt.device));
    const bool learned = c->learned;
    c->learned = false;                 // tunes the plain seed
    // the cone search's leaf size, seed and form are tuned on the cone form
    const int rs0 = c->opt.round_search;
    c->opt.round_search = 0;
    float best = 0.f;
    int bb = 64;
    float bs = c->seed_scale;
    for (int b : {16, 32, 64})
        for (float sc : {2.f, 3.f, 4.f}) {
            if ((rc = layout_clouds(c, b))) break;
            c->seed_scale = sc;
            c->no_grouping = false;
            if ((rc = build(c))) break;
            const float t = c->times[0] + c->times[1];
            if (best == 0.f || t < best) {
                best = t;
                bb = b;
                bs = sc;
            }
        }
    // then the form of the cone search for that pair (one launch, or resumed
    // in place or compacted after the first pass)
    const int split0 = c->opt.search_split;
    int bsplit = split0;
    if (!rc) rc = layout_clouds(c, bb);
    if (!rc) {
        c->seed_scale = bs;
        for (int sp : {1, 2, 3}) {
            c->opt.search_split = sp;
            c->no_grouping = false;
            if ((rc = build(c))) break;
            const float t = c->times[0] + c->times[1];
            if (t < best) {
                best = t;
                bsplit = sp;
            }
        }
    }
    c->opt.search_split = rc ? split0 : bsplit;
    // then the search itself: the round-1 cone search or the per-round
    // component search, each timed over a whole run (their forests are the
    // same).  Kept as the caller set it with a learned model (a model
    // belongs to one search) or in the 2.5D mode (cone search only).
    int brs = rs0;
    float bscale = bs;
    if (!rc && !learned && c->cdim == c->dim) {
        float tbest = 0.f;
        int bbest = bb;
        for (int rs : {0, 1}) {
            c->opt.round_search = rs;
            c->seed_scale = rs == 1 && c->dim == 3 ? kCompSeedScale3D : bs;
            // the component search at its own leaf size, 64 points
            const int bk = rs == 1 ? 64 : bb;
            c->no_grouping = false;
            if ((rc = layout_clouds(c, bk)) || (rc = build(c)) || (rc = run_rounds(c))) break;
            const float t = c->times[0] + c->times[1] + c->times[2];
            if (tbest == 0.f || t < tbest) {
                tbest = t;
                brs = rs;
                bscale = c->seed_scale;
                bbest = bk;
            }
        }
        bb = bbest;
    }
    c->opt.round_search = brs;
    bs = bscale;
    c->learned = learned;
    if (rc) return rc;
    c->opt.bucket_size = bb;
    c->seed_scale = bs;
    if ((rc = layout_clouds(c, bb))) return rc;
    c->built = false;
    c->finished = false;
    c->rounds = 0;
    c->edges = 0;
    c->trace.clear();
    if (best_bucket) *best_bucket = bb;
    if (best_scale) *best_scale = bs;
    if (best_ms) *best_ms = best;
    return EMST_OK;
}

int emst_get_cloud_results(emst_ctx* c, double* weight, int32_t* edges, int32_t* last_round)
{
    int rc = need_points(c);
    if (rc) return rc;
    if (!c->built) return fail(c, EMST_E_STATE, "no run yet");
    HIPCHK(c, hipSetDevice(c->opt.device));
    const int ncl = c->ncl;
    if ((rc = ensure(c, c->cloud_w, (size_t)ncl)) || (rc = ensure(c, c->cloud_e, (size_t)ncl)) ||
        (rc = ensure(c, c->cloud_r, (size_t)ncl)))
        return rc;
    hipStream_t s = c->stream;
    launch_cloud_results(c->xyz, c->dim, c->dcl.p, ncl, c->eu.p, c->ev_.p, c->eround.p, c->cloud_w.p,
                         c->cloud_e.p, c->cloud_r.p, s);
    LAUNCHCHK(c);
    if (weight) HIPCHK(c, hipMemcpyAsync(weight, c->cloud_w.p, sizeof(double) * ncl, hipMemcpyDeviceToHost, s));
    if (edges) HIPCHK(c, hipMemcpyAsync(edges, c->cloud_e.p, sizeof(int32_t) * ncl, hipMemcpyDeviceToHost, s));
    if (last_round)
        HIPCHK(c, hipMemcpyAsync(last_round, c->cloud_r.p, sizeof(int32_t) * ncl, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return EMST_OK;
}

int emst_run(emst_ctx* c)
{
    int rc = need_points(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->opt.device));
    if (!c->built && one_launch_on(c)) return run_one_launch(c);
    rc = ensure_built(c);
    if (rc) return rc;
    return run_rounds(c);
}

int emst_step(emst_ctx* c, int32_t* finished)
{
    if (!finished) return fail(c, EMST_E_ARG, "finished must not be NULL");
    int rc = need_points(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->opt.device));
    rc = ensure_built(c);
    if (rc) return rc;
    return round_once(c, finished);
}

int emst_get_edges(emst_ctx* c, int32_t* u, int32_t* v, float* d2, int64_t* m)
{
    if (!m) return fail(c, EMST_E_ARG, "m must not be NULL");
    int rc = need_points(c);
    if (rc) return rc;
    const int64_t cnt = c->built ? (int64_t)c->edges : 0;
    if (!u && !v && !d2) { *m = cnt; return EMST_OK; }
    if (*m < cnt) { *m = cnt; return fail(c, EMST_E_ARG, "edge arrays too small"); }
    *m = cnt;
    if (cnt == 0) return EMST_OK;
    HIPCHK(c, hipSetDevice(c->opt.device));
    hipStream_t s = c->stream;
    // slots are indexed by absorbed root; empty slots sort last
    launch_edge_keys(c->eu.p, c->ev_.p, c->n, c->keys0.p, s);
    radix_sort(c->keys0.p, c->vals0.p, c->keys1.p, c->vals1.p, c->n, 64, true, c->hist.p, c->scan_tmp.p, s);
    // staging: the sort's second key buffer (free once the sort has landed in
    // keys0) and the research list (rebuilt in every round)
    int32_t* du = reinterpret_cast<int32_t*>(c->keys1.p);
    int32_t* dv = c->list.p;
    float* dd = reinterpret_cast<float*>(c->vals1.p);
    launch_edges_out(c->keys0.p, c->vals0.p, c->ed2.p, cnt, du, dv, dd, s);
    LAUNCHCHK(c);
    if (u) HIPCHK(c, hipMemcpyAsync(u, du, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, s));
    if (v) HIPCHK(c, hipMemcpyAsync(v, dv, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, s));
    if (d2) HIPCHK(c, hipMemcpyAsync(d2, dd, sizeof(float) * cnt, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return EMST_OK;
}

int emst_get_weight(emst_ctx* c, double* w, int64_t* size)
{
    if (!w) return fail(c, EMST_E_ARG, "weight must not be NULL");
    int rc = need_points(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->opt.device));
    const int64_t cnt = c->built ? (int64_t)c->edges : 0;
    if (size) *size = cnt;
    return weight_of(c, cnt > 0 ? c->n : 0, w);
}

int emst_weight_noalloc(const float* xyz, int32_t dim, const int32_t* eu, const int32_t* ev, int64_t m,
                        int32_t device, void* stream, double* weight)
{
    if (!weight) { g_create_error = "weight must not be NULL"; return EMST_E_ARG; }
    *weight = 0;
    if (dim != 2 && dim != 3) { g_create_error = "dim must be 2 or 3"; return EMST_E_ARG; }
    if (m < 0) { g_create_error = "m < 0"; return EMST_E_ARG; }
    if (m >= ((int64_t)1 << 31)) { g_create_error = "m must be below 2^31"; return EMST_E_LIMIT; }
    if (m > 0 && (!xyz || !eu || !ev)) { g_create_error = "NULL device array"; return EMST_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        g_create_error = "no HIP device";
        return EMST_E_NODEVICE;
    }
    if (device < 0 || device >= ndev || device >= kMaxDevices) {
        g_create_error = "device ordinal out of range";
        return EMST_E_ARG;
    }
    hipError_t e = hipSetDevice(device);
    double* part = nullptr;
    if (e == hipSuccess) e = scratch_on(device, &part);
    if (e != hipSuccess) {
        g_create_error = std::string("scratch set-up: ") + hipGetErrorString(e);
        return EMST_E_HIP;
    }
    if (m == 0) return EMST_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    launch_weight(xyz, dim, eu, ev, m, part, kWeightBlocks, s);
    double h[kWeightBlocks];
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        g_create_error = std::string("weight: ") + hipGetErrorString(e);
        return EMST_E_HIP;
    }
    double sum = 0;
    for (double x : h) sum += x;
    *weight = sum;
    return EMST_OK;
}

int emst_get_labels(emst_ctx* c, int32_t* labels)
{
    if (!labels) return fail(c, EMST_E_ARG, "labels must not be NULL");
    int rc = ensure_built(c);
    if (rc) return rc;
    if (c->n == 0) return EMST_OK;
    launch_labels_out(c->n, c->label.p, c->pts.p, c->inv.p, c->list.p, c->stream);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(labels, c->list.p, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMST_OK;
}

int emst_get_rescaled(emst_ctx* c, float* xyz, double* unit_x)
{
    int rc = ensure_built(c);
    if (rc) return rc;
    if (xyz && c->n > 0)
        HIPCHK(c, hipMemcpyAsync(xyz, c->resc.p, sizeof(float) * c->n * c->dim, hipMemcpyDeviceToHost, c->stream));
    if (unit_x) HIPCHK(c, hipMemcpyAsync(unit_x, &c->P->unit_x, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMST_OK;
}

int emst_get_trace(emst_ctx* c, emst_round_trace* rows, int32_t* nrows)
{
    if (!c || !nrows) return EMST_E_ARG;
    const int32_t cnt = (int32_t)c->trace.size();
    if (!rows) { *nrows = cnt; return EMST_OK; }
    const int32_t k = std::min(*nrows, cnt);
    for (int32_t i = 0; i < k; ++i) rows[i] = c->trace[i];
    *nrows = k;
    return EMST_OK;
}

int emst_get_candidates(emst_ctx* c, int32_t* cand, float* cd2, int32_t* k)
{
    int rc = ensure_built(c);
    if (rc) return rc;
    if (c->opt.round_search == 1) return fail(c, EMST_E_STATE, "the per-round component search keeps no cone candidates");
    if (k) *k = c->K;
    if (!cand && !cd2) return EMST_OK;
    if (!cand || !cd2) return fail(c, EMST_E_ARG, "cand and cand_d2 are both required");
    if (c->n == 0) return EMST_OK;
    // staging in keys1/vals1 would be too small for n*K: use cand_pos-sized scratch
    DevBuf<int32_t> tc;
    DevBuf<float> td;
    if ((rc = ensure(c, tc, (size_t)(c->n * c->K))) != EMST_OK) return rc;
    if ((rc = ensure(c, td, (size_t)(c->n * c->K))) != EMST_OK) { release(tc); return rc; }
    launch_cands_out(c->K, c->n, c->cand_key.p, c->inv.p, tc.p, td.p, c->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cand, tc.p, sizeof(int32_t) * c->n * c->K, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cd2, td.p, sizeof(float) * c->n * c->K, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    release(tc);
    release(td);
    if (e != hipSuccess) return fail(c, EMST_E_HIP, std::string("candidates: ") + hipGetErrorString(e));
    return EMST_OK;
}

int emst_get_cell_grid(emst_ctx* c, int32_t R, int32_t* ext_dual, int32_t* cell_of, int64_t* cell_start,
                       int32_t* cell_pts)
{
    if (!ext_dual) return fail(c, EMST_E_ARG, "ext_dual must not be NULL");
    if (R <= 0) return fail(c, EMST_E_ARG, "level radius must be positive");
    int rc = ensure_built(c);
    if (rc) return rc;
    if (c->ncl != 1) return fail(c, EMST_E_STATE, "the reference cell grid is defined for one cloud");
    Params hp{};
    HIPCHK(c, hipMemcpyAsync(&hp, c->P, sizeof(Params), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // extents (SolutionEMST.cu:420-430) and ViewGridQuadMD recurrences (ViewGrid.h:251-265);
    // the rescaled minimum is 0 (the box corner maps to 0 exactly).
    int32_t geo[16] = {0};
    for (int k = 0; k < 3; ++k) {
        int32_t ext = 1, pc = 0, PC = 0, ed = 1;
        if (k < c->cdim) {
            ext = (int32_t)std::ceil((double)hp.newmax[k] - 0.0) + 1;
            pc = ext / 2;
            PC = (pc + (R - 1)) / R;
            const int32_t eb = (PC + 1) + ((ext - pc) + (R - 1)) / R;
            const int32_t PCD = PC / 2;
            ed = (PCD + 1) + ((eb - PC) + 1) / 2;
        }
        geo[k] = ed;
        geo[3 + k] = PC;
        geo[6 + k] = pc;
    }
    geo[9] = R;
    geo[10] = c->cdim;
    for (int k = 0; k < 3; ++k) ext_dual[k] = geo[k];
    if (!cell_of && !cell_start && !cell_pts) return EMST_OK;
    const int64_t ncell = (int64_t)geo[0] * geo[1] * geo[2];
    const int64_t n = c->n;
    DevBuf<int32_t> dcell;
    DevBuf<uint32_t> dcount, dstart, tmp;
    auto cleanup = [&]() { release(dcell); release(dcount); release(dstart); release(tmp); };
    if ((rc = ensure(c, dcell, (size_t)(3 * n))) || (rc = ensure(c, dcount, (size_t)ncell + 1)) ||
        (rc = ensure(c, dstart, (size_t)ncell + 1)) || (rc = ensure(c, tmp, scan_tmp_words(ncell + 1) + 16))) {
        cleanup();
        return rc;
    }
    hipStream_t s = c->stream;
    hipError_t e = hipMemcpyAsync(c->geo.p, geo, sizeof(geo), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(dcount.p, 0, sizeof(uint32_t) * (ncell + 1), s);
    if (e == hipSuccess) {
        launch_cell_grid(c->resc.p, n, c->cdim, c->dim, c->geo.p, dcell.p, c->keys0.p, dcount.p, s);
        exclusive_scan_u32(dcount.p, dstart.p, ncell + 1, tmp.p, s);
        int bits = 8;
        while (bits < 64 && ((uint64_t)ncell >> bits) != 0) bits += 8;
        radix_sort(c->keys0.p, c->vals0.p, c->keys1.p, c->vals1.p, n, bits, true, c->hist.p, c->scan_tmp.p, s);
        e = hipGetLastError();
    }
    std::vector<uint32_t> hstart;
    if (e == hipSuccess && cell_of) e = hipMemcpyAsync(cell_of, dcell.p, sizeof(int32_t) * 3 * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && cell_start) {
        hstart.resize(ncell + 1);
        e = hipMemcpyAsync(hstart.data(), dstart.p, sizeof(uint32_t) * (ncell + 1), hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess && cell_pts) e = hipMemcpyAsync(cell_pts, c->vals0.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return fail(c, EMST_E_HIP, std::string("cell grid: ") + hipGetErrorString(e));
    if (cell_start)
        for (int64_t i = 0; i <= ncell; ++i) cell_start[i] = hstart[i];
    // the count of points inside the grid closes the CSR (outside points sort last)
    return EMST_OK;
}

int emst_get_shape(emst_ctx* c, int64_t* n, int32_t* dim)
{
    int rc = need_points(c);
    if (rc) return rc;
    if (n) *n = c->n;
    if (dim) *dim = c->dim;
    return EMST_OK;
}

int emst_get_options(emst_ctx* c, emst_options* o)
{
    if (!c || !o) return EMST_E_ARG;
    *o = c->opt;
    o->seed_scale = c->seed_scale;
    return EMST_OK;
}

int emst_get_times(emst_ctx* c, float* times, int32_t* n_times)
{
    if (!c || !times || !n_times) return EMST_E_ARG;
    const int32_t k = std::min<int32_t>(*n_times, 6);
    for (int i = 0; i < k; ++i) times[i] = i < 4 ? c->times[i] : (i == 4 ? c->comp_search_ms : c->knn_ms);
    *n_times = k;
    return EMST_OK;
}

int emst_get_round_ends(emst_ctx* c, float* ms, int32_t* n)
{
    if (!c || !n) return EMST_E_ARG;
    const int32_t have = (int32_t)c->round_end_ms.size();
    if (!ms) {
        *n = have;
        return EMST_OK;
    }
    const int32_t k = std::min(*n, have);
    for (int32_t i = 0; i < k; ++i) ms[i] = c->round_end_ms[i];
    *n = k;
    return EMST_OK;
}

int emst_set_radius_model(emst_ctx* c, const float* model)
{
    if (!c) return EMST_E_ARG;
    HIPCHK(c, hipSetDevice(c->opt.device));
    if (!model) {
        c->learned = false;
        c->built = false;
        return EMST_OK;
    }
    if (c->opt.cone_dim == 2 && c->opt.dim == 3)
        return fail(c, EMST_E_ARG, "the learned seed radius is not defined for the 2.5D mode");
    int rc = ensure(c, c->model, EMST_MLP_PARAMS);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->model.p, model, sizeof(float) * EMST_MLP_PARAMS, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->learned = true;
    c->built = false;       // the next run searches with the new seeds
    return EMST_OK;
}

int emst_get_radius_samples(emst_ctx* c, float* feat, float* target, int64_t* m)
{
    if (!m) return fail(c, EMST_E_ARG, "m must not be NULL");
    int rc = ensure_built(c);
    if (rc) return rc;
    if (c->ncl != 1) return fail(c, EMST_E_STATE, "radius samples come from a single-cloud context");
    if (c->cdim != c->dim) return fail(c, EMST_E_STATE, "radius samples come from full-dimensional cones");
    if (c->opt.round_search == 1) {
        if (c->opt.log_samples != 1)
            return fail(c, EMST_E_STATE, "the component search logs samples with options.log_samples = 1");
        Counters h{};
        if ((rc = read_counters(c, &h))) return rc;
        const int64_t got = std::min<int64_t>((int64_t)h.nsamples, kSampleCap);
        if (!feat && !target) { *m = got; return EMST_OK; }
        if (*m < got) { *m = got; return fail(c, EMST_E_ARG, "sample arrays too small"); }
        *m = got;
        if (got == 0) return EMST_OK;
        std::vector<float> h9((size_t)got * 9);
        HIPCHK(c, hipMemcpyAsync(h9.data(), c->samples.p, sizeof(float) * h9.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < got; ++i) {
            if (feat) std::memcpy(feat + i * 8, &h9[(size_t)i * 9], sizeof(float) * 8);
            if (target) target[i] = h9[(size_t)i * 9 + 8];
        }
        return EMST_OK;
    }
    const int64_t n = c->n;
    if (!feat && !target) { *m = n; return EMST_OK; }
    if (*m < n) { *m = n; return fail(c, EMST_E_ARG, "sample arrays too small"); }
    *m = n;
    if (n == 0) return EMST_OK;
    if ((rc = ensure(c, c->feat, (size_t)n * 8)) || (rc = ensure(c, c->pred, (size_t)n)) ||
        (rc = ensure(c, c->ls, (size_t)n)))
        return rc;
    hipStream_t s = c->stream;
    launch_radius_features(c->dim, c->pts.p, c->nodes.p, c->L, leaf_table(c), n, c->feat.p, c->ls.p, s);
    launch_radius_targets(c->K, n, c->cand_key.p, c->ls.p, c->pred.p, s);
    LAUNCHCHK(c);
    if (feat) HIPCHK(c, hipMemcpyAsync(feat, c->feat.p, sizeof(float) * n * 8, hipMemcpyDeviceToHost, s));
    if (target) HIPCHK(c, hipMemcpyAsync(target, c->pred.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return EMST_OK;
}

int emst_get_search_stats(emst_ctx* c, int64_t* stats, int32_t* n_stats)
{
    if (!c || !stats || !n_stats) return EMST_E_ARG;
    const int32_t k = std::min<int32_t>(*n_stats, 18);
    for (int i = 0; i < k; ++i) stats[i] = c->stats[i];
    *n_stats = k;
    return EMST_OK;
}

const char* emst_version(void) { return "emst_amd 0.1 (gfx950)"; }

}  // extern "C"
