// emst_internal.h -- shared types of the EMST engine (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emst {

constexpr uint64_t kEmptyKey = ~0ull;   // empty cone / no candidate
constexpr int kWave = 64;

// Device-side scalar parameters of one point set.
struct Params {
    float pmin[3];          // bounding box of the input (A1)
    float pmax[3];
    double unit_x;          // rescale factor (A2), double as computed
    float unit_xf;          // (float)unit_x, the factor applied per coordinate
    float newmax[3];        // rescaled extremes (newmin is 0 by monotonicity)
    float qscale;           // Morton quantisation scale
    uint32_t nonfinite;     // set when a coordinate is NaN or Inf
    uint32_t pad[3];
};

constexpr int kWorkSlots = 64;
enum {
    kWorkCompNodes, kWorkCompLeaves, kWorkCompPairs, kWorkKnnNodes, kWorkKnnLeaves, kWorkKnnPairs,
    kWorkCompWalks, kWorkKnnWalks, kWorkKinds
};

// Round counters, one device struct per context.
struct Counters {
    uint32_t edges;         // total forest edges
    uint32_t research;      // points queued for a repeated cone search
    uint32_t zero_or_empty; // points holding a zero-distance or no candidate
    uint32_t unflat;        // flatten stopped at its step cap (run it again)
    // cone-search counters of the last full search: packets (waves), nodes
    // tested, leaves entered, leaf points read per packet, lane-point pairs
    unsigned long long packets, nodes, leaves, leaf_points, pairs;
    // coincident-point groups (set once per build, not reset by the rounds)
    uint32_t dup_runs;      // runs of >= 2 equal Morton keys
    uint32_t dup_points;    // points with an identical twin (group members + representatives)
    uint32_t dup_overflow;  // a run too long to split into groups: grouping off
    uint32_t grouped;       // grouping on (dup_points > 0 and no overflow)
    uint32_t underflow;     // two distinct points at float distance 0 seen: grouping must be off
    uint32_t dup_groups;    // number of groups
    uint32_t dup_split;     // set by k_dup_m2 when some group spans >= 2 components
    uint32_t cycles_cut;    // hooks withheld by the cycle rule (2.5D mode), over the run
    // round control, on the device so that rounds can run back to back
    // (captured in a graph) and stop by themselves
    uint32_t round;         // rounds run so far
    uint32_t done;          // 0 running, 1 finished, 2 paused: flatten needs another launch
    uint32_t edges_prev;    // edges before the current round
    uint32_t target;        // sum over clouds of (points - 1)
    uint32_t max_rounds;
    uint32_t dup_settled;   // no coincident group spans two components any more
    uint32_t dup_big;       // runs longer than one wave handles, queued for a grid-wide pass
    uint32_t nroots;        // roots listed by k_cycle_init this round (2.5D cycle rule)
    uint32_t nsamples;      // seed samples logged by the component search over the run
    unsigned long long comp_queries;    // queries over the component search's dispatches, over the run
    // work of the component search's dispatches over the run (node tests,
    // one per node a packet examines; leaf visits; point pairs evaluated),
    // then the same for the neighbour lists at build (k_knn_first), then the
    // passes of each that walked the tree per query because their gathering
    // outgrew its lists or had no finite bound; each
    // spread over kWorkSlots words (a packet adds to slot blockIdx % kWorkSlots,
    // so that the adds do not queue on one address), summed by the host
    unsigned long long work[kWorkKinds][kWorkSlots];
    // a build issued without host checks (one launch) found a NaN or Inf
    // coordinate or a cloud with more leaves than its slots: the rounds do
    // nothing and the host redoes or refuses the run
    uint32_t invalid;
    // queries of this round's component search handed to the per-query pass
    // (k_comp_search_deferred), entries of the defer list
    uint32_t deferred;
    // round 1's targets were written before the rounds (k_knn_first): round 1
    // then needs no keep, prep, node labels or second key, and its hook takes
    // every point with a target as its singleton component's winner
    uint32_t knn_round1;
};

// Coincident-point groups.  Points with identical rescaled coordinates form a
// group whose representative is its smallest input id (its first sorted
// position, the sort being stable).  rep[p] = representative's sorted
// position, or -1 for a point without an identical twin.  While grouping is
// on, only representatives and singletons search, once, and no point takes a
// candidate at its own location; k_dup_fill then gives every group point its
// cones: the '-' cones are the representative's, the '+' cones either the
// zero-distance candidate of the group (while the group spans >= 2
// components) or the representative's own '+' cones kept in bkey/bpos.
struct DupState {
    const int32_t* rep;     // nullptr: grouping off
    const int32_t* list;    // the group points (members and representatives)
    const int32_t* gslot;   // per representative position: its group index
    uint32_t* m2;           // per group: smallest member id whose label differs from the representative's
    uint64_t* bkey;         // per group: the representative's '+' cones, [D * 2^(D-1)][ngroups]
    int32_t* bpos;
    int64_t npoints;        // entries of list
    int64_t ngroups;
};

// Cycle rule of the 2.5D mode (wedges over the xy plane, 3D distances).
// There a component's chosen edge need not be the other side's smallest, and
// the choices can close a cycle of components, of length >= 3 or of length 2
// with two different edges; hooking along it would leave a cyclic label
// chain.  Per round, over the component roots: j0 = the root that the chosen
// edge leads to (the root itself when it has no outgoing candidate or chose
// the same edge as that root: a mutual pair, which the hook resolves); the
// roots with another target are listed (roots, Counters::nroots); pointer
// doubling over the list (j, kh/kl ping-pong, S steps with 2^S > listed
// roots, at most `steps`) gives the largest chosen edge (kh = d2 bits << 32 |
// min id, kl = max id) on the path of 2^S hops, which for a root on a cycle
// is the cycle's largest; onc marks the roots on cycles, cut the root whose
// own edge is its cycle's largest: that root does not hook this round.
// j0 == nullptr: off.
struct CycleState {
    int32_t* roots;         // listed roots (n)
    int32_t* j0;
    int32_t* j[2];
    uint64_t* kh[2];
    uint32_t* kl[2];
    uint8_t* onc;
    uint8_t* cut;
    const int32_t* inv;     // global id -> sorted position
    int steps;
};

// One point cloud of a context (a context holds one cloud, or a batch of
// independent clouds of one dimension run together, packed back to back in
// one input array).  Cloud c owns the global ids and sorted positions
// [pos, pos + n) (ids keep input order inside the cloud, so every tie breaks
// as in a run of the cloud alone; its input coordinates start at float
// pos * dim) and its own implicit binary hierarchy of L leaf slots (a power
// of two) at nodes[2 * lbase + i], 1 <= i < 2L.  Its nl leaves are the global
// leaves [lfirst, lfirst + nl) of the LeafTable, in slots 0 .. nl - 1; the
// slots past them are empty.
struct Cloud {
    int64_t pos;
    int64_t n;
    int64_t lbase;
    int64_t L;
    int64_t lfirst;
    int64_t nl;
};

// Leaves of the search hierarchy: the maximal cells of the binary prefix tree
// of the sorted keys (Morton prefixes) that hold at most `bucket` points; a
// run of more than `bucket` equal keys, which no key bit splits, is cut every
// `bucket` points from its start.  Numbered in sorted order, cloud by cloud:
// leaf g holds the sorted positions [start[g], start[g + 1]); of[p] = the
// leaf of position p.  Aligned cells keep the leaf boxes tight (a fixed
// number of consecutive positions would straddle the jumps of the curve).
struct LeafTable {
    const int32_t* start;   // leaves + 1
    const int32_t* of;      // n
};

// Search-hierarchy node: tight box of a contiguous range of Morton-sorted points.
struct alignas(32) Node {
    float lo[3];
    float hi[3];
    // unused by the cone search; the component search keeps here the bits of
    // the node's smallest point id (pad[0], per build) and of its component
    // label (pad[1], per round: the common label of its points, -1 mixed,
    // -3 empty)
    float pad[2];
};

// ---- launch wrappers (emst_kernels.hip); all asynchronous on `s` ----
// Per cloud: bounding box (A1), rescale parameters (A2); per point: rescaled
// coordinates and the sort key (cloud index above mbits-per-axis Morton bits;
// no cloud bits for a single cloud).  N = total points, P[ncl].
void launch_bbox(const float* xyz, const Cloud* cl, int ncl, int64_t N, int dim, Params* P, hipStream_t s);
void launch_params(Params* P, const Cloud* cl, int ncl, int dim, int mbits, hipStream_t s);
void launch_rescale_morton(const float* xyz, const Cloud* cl, int ncl, int64_t N, int dim, int mbits,
                           const Params* P, float* resc, uint64_t* keys, hipStream_t s);
// LSD radix sort of (keys, vals) by bits [0, bits); result lands in (k0, v0).
// vals_identity: values start as 0..n-1 (v0 is not read).
void radix_sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int bits,
                bool vals_identity, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s);
size_t radix_hist_words(int64_t n);
size_t scan_tmp_words(int64_t n);
void launch_gather_sorted(const float* resc, const uint32_t* perm, int64_t n, int dim, float4* pts,
                          int32_t* inv, uint32_t* ids, hipStream_t s);
// Coincident-point groups from the sorted keys and points: fills rep[] (all
// n entries), list[], gslot[] (representatives) and C->dup_* / C->grouped.
// runs: n words of scratch.  The dup counters must be zero on entry.
void launch_dup_groups(const uint64_t* keys, const float4* pts, int64_t n, int32_t* rep,
                       int32_t* runs, int32_t* list, int32_t* gslot, Counters* C, hipStream_t s);
// Leaf boxes leave out group members (rep == nullptr: grouping off).
// nleaves = sum of the clouds' L, Lmax = largest L.
void launch_build_tree(const float4* pts, const Cloud* cl, int ncl, int64_t nleaves, int64_t Lmax, LeafTable lt,
                       Node* nodes, const int32_t* rep, hipStream_t s);
// The leaves of the sorted keys (see LeafTable): mark and idx are n words of
// scratch, tmp scan_tmp_words(n); start (n + 1 words at most) and of (n) are
// written, and cloud_lf[c] = the first leaf of cloud c, cloud_lf[ncl] = the
// number of leaves.
// layout: each cloud's slot count L is fixed beforehand (a bound); the
// leaves found are written into the cloud table on the device, and
// C->invalid set when a cloud has more leaves than slots or a NaN or Inf
// coordinate (P).  Without layout the host lays the hierarchy out from
// cloud_lf.
void launch_leaves(const uint64_t* keys, int64_t n, Cloud* cl, int ncl, int bucket, uint32_t* mark,
                   uint32_t* idx, uint32_t* tmp, int32_t* start, int32_t* of, int32_t* cloud_lf, const Params* P,
                   Counters* C, bool layout, hipStream_t s);
void launch_zero_u32(uint32_t* p, int64_t n, hipStream_t s);
// Cone search over queries [0,n) (qlist == nullptr) or qlist[0..nq).
// check_labels: skip zero-distance points of the query's own component.
// collect_stats: add the traversal counters to C.
// Scratch of a split cone search (pass 0 over all queries, the rest over the
// queries left open, compacted): P = packets of 64 queries.
struct SplitScratch {
    uint64_t* open_mask;    // P
    uint32_t* cnt;          // P + 1
    uint32_t* off;          // P + 1
    uint32_t* tmp;          // scan_tmp_words(P + 1)
    int32_t* list;          // n
    int32_t* list2;         // n (component search: the queries left open after pass 0)
    uint64_t* defer;        // n (component search: queries handed to the per-query pass)
    int after;              // the first launch runs passes 0..after
    double compact_frac;    // resume from a compacted list when at most this share of the queries is open
};
// dim: point dimension (distances); cdim: cone dimension (cdim = 2 with dim =
// 3 is the 2.5D mode: wedges over the xy plane).  split (nullptr: one launch
// over every pass) applies to full searches (qlist == nullptr).
void launch_cone_search(int dim, int cdim, const float4* pts, const Node* nodes, const Cloud* cl, int ncl, int bucket,
                        LeafTable lt, int64_t n, const int32_t* qlist, int64_t nq, const int32_t* label,
                        bool check_labels, const int32_t* inv, const float* seed_r, float seed_scale,
                        uint64_t* cand_key, int32_t* cand_pos, int32_t* revive, int32_t* first_corr,
                        uint64_t* first_key, Counters* C, bool collect_stats, DupState dup,
                        const SplitScratch* split, hipStream_t s);
// Phase 0 of a round of the per-round component search (the alternative to
// the round-1 cone search): node component labels (into Node::pad[1]; visit:
// the arrival words of launch_node_min_ids), kept targets, and a search for
// the points whose target joined their component; writes corr / corr_key and
// the component minima (comp1).  sc: open_mask, cnt, off, tmp, list and
// list2; pass 0 runs over every searching point, the later passes over those
// it left open, so that they see the component minima pass 0 found.
// Learned seed of the component search: per-component size and box, kept at
// the roots and merged after every round; features of the searching points
// (list order), the model's output and the seed radius per sorted position.
// features_on: sizes, boxes and features are maintained (a model, or logging).
struct CompSeed {
    bool features_on;
    int32_t* csize;         // n
    uint32_t* cbox;         // 6n: f2ord of lo x, y, z, hi x, y, z
    float* X;               // 8n
    float* LS;              // n: log of the leaf spacing
    float* Y;               // n
    float* R;               // n
    const float* model;     // nullptr: plain seeds
    float* samples;         // 9 floats per sample (8 features, target), or nullptr
    int64_t cap;            // samples that fit
    int64_t per_round;      // pairs logged per round, at most (a stride over the searching points)
    uint32_t* nq_keep;      // 2 words: the counts the sample log and the model run on
};
// Neighbour lists of the component search: every point's K nearest other
// points of its cloud under (d2, id), as positions and keys, list entry k of
// point p at k * n + p (position -1, key all ones: the cloud has no more
// points), and per point the entry its current target came from (K: the list
// is used up).  K = knn_len(dim): longer lists in 3D, where they answer the
// searches of the early rounds (clustered 3D 10^7: 6 entries 14.6 ms per
// EMST, 8 13.7, 10 13.5, 12 14.9); 2D is fastest at 6 (12.3 ms; 8: 12.4,
// 10: 12.6).
constexpr int kKnn2 = 6;
constexpr int kKnn3 = 10;
__host__ __device__ constexpr int knn_len(int dim) { return dim == 3 ? kKnn3 : kKnn2; }
struct NbrList {
    int32_t* pos;           // K * n
    uint64_t* key;          // K * n
    uint8_t* cur;           // n
};
void launch_comp_round(int dim, int64_t n, const float4* pts, const uint32_t* ids, Node* nodes,
                       unsigned long long* visit,
                       const Cloud* cl, int ncl, int64_t nleaves, int bucket, LeafTable lt, const int32_t* label, int32_t* corr, uint64_t* corr_key, uint64_t* comp1, const int32_t* inv,
                       float seed_scale, const SplitScratch& sc, const CompSeed& cs, const NbrList& nb,
                       const int32_t* eu, const int32_t* eround, Counters* C, hipStream_t s,
                       hipEvent_t* ev = nullptr);
// The round-1 targets of the component search and the neighbour lists: one
// traversal per 64-point packet (after launch_round_init and
// launch_node_min_ids).  Writes corr / corr_key (every component is one
// point, so the nearest neighbour is the round's exact target) and nb.
void launch_knn_first(int dim, int64_t n, const float4* pts, const Node* nodes, const Cloud* cl, int ncl,
                      int bucket, LeafTable lt, const int32_t* inv, int32_t* corr, uint64_t* corr_key,
                      const NbrList& nb, Counters* C, bool collect, hipStream_t s);
// (ev: 4 events recorded around the two k_comp_search dispatches, or nullptr)
// Sizes and boxes of the single-point components (the start of a run).
void launch_comp_seed_init(int64_t n, const float4* pts, int32_t* csize, uint32_t* cbox, hipStream_t s);
// MLP inference over X[0..*m_dev) (the comp search's device-side count).
void launch_mlp_infer_dev(const float* model, const float* X, const uint32_t* m_dev, int64_t m_max, float* Y,
                          hipStream_t s);
// Smallest point id under every node of the hierarchy, into Node::pad[0]
// (0xffffffff for a node without points).  visit:
// 2 * nleaves words of arrival state, zero before the first bottom-up launch
// over this hierarchy (k_nmin_tree, k_nlab_tree) and kept between them.
void launch_node_min_ids(const float4* pts, const Cloud* cl, int ncl, int64_t nleaves, LeafTable lt, Node* nodes,
                         unsigned long long* visit, hipStream_t s);
// Cones of every group point for the current labels (dim: the cone dimension) (after the grouped
// cone search, and at the start of each round while a group may still span
// several components).  Sets C->dup_split when one does.
void launch_dup_fill(int dim, int64_t n, const int32_t* label, const float4* pts, const int32_t* inv,
                     uint64_t* cand_key, int32_t* cand_pos, int32_t* revive, int32_t* first_corr,
                     uint64_t* first_key, Counters* C, DupState dup, hipStream_t s);
void launch_round_init(int64_t n, int32_t* label, int32_t* label2, int32_t* eu, int32_t* corr,
                       uint64_t* corr_key, uint64_t* comp1, uint32_t* comp2, Counters* C, uint32_t target,
                       uint32_t max_rounds, hipStream_t s);
// Phase of a round: 0 per-point target + component min (first key), 1 second
// key, 2 winners and hook, 3 flatten.
// Ranks each point's cones by key for k_round_min (all points, or the qlist
// points); C != nullptr: skipped once the rounds are done or groups settled.
// dim: the cone dimension.
void launch_rank_cones(int dim, int64_t n, const uint64_t* cand_key, const int32_t* cand_pos,
                       const int32_t* qlist, int64_t nq, uint64_t* rkey, int32_t* rpos, uint8_t* cursor,
                       const Counters* C, hipStream_t s);
// rkey / rpos: the ranked candidates of k_rank_cones (point-major, K per point)
void launch_round_phase(int phase, int K, int64_t n, const uint32_t* ids, const uint64_t* rkey,
                        const int32_t* rpos, uint8_t* cursor, int32_t* label,
                        int32_t* label2, int32_t* corr,
                        uint64_t* corr_key, uint64_t* comp1, uint32_t* comp2,
                        Counters* C, int32_t* eu, int32_t* ev, float* ed2, int32_t* eround,
                        const CycleState& cy, hipStream_t s);
// Round control: begin (termination test, round += 1) and end (no progress
// -> finished; flatten unfinished -> paused; groups settled).
void launch_round_begin(Counters* C, hipStream_t s);
void launch_round_end(Counters* C, bool grouped, hipStream_t s, unsigned long long* stamps = nullptr,
                      uint32_t cap = 0);
// the device's constant-rate clock (hipDeviceAttributeWallClockRate) into *at
void launch_stamp(unsigned long long* at, hipStream_t s);
// Queue the points holding a zero-distance or no candidate.
void launch_flag_research(int K, int64_t n, const uint64_t* cand_key, int32_t* list,
                          Counters* C, hipStream_t s);
void launch_weight(const float* xyz, int dim, const int32_t* eu, const int32_t* ev, int64_t m,
                   double* partial, int nblocks, hipStream_t s);
// Per cloud: FP64 weight of its forest edges (input coordinates), its edge
// count and the last round that added one of them.
void launch_cloud_results(const float* xyz, int dim, const Cloud* cl, int ncl, const int32_t* eu,
                          const int32_t* ev, const int32_t* eround, double* weight, int32_t* edges,
                          int32_t* last_round, hipStream_t s);
void launch_labels_out(int64_t n, const int32_t* label, const float4* pts, const int32_t* inv,
                       int32_t* out, hipStream_t s);
void launch_cands_out(int K, int64_t n, const uint64_t* cand_key, const int32_t* inv,
                      int32_t* cand, float* cd2, hipStream_t s);
void launch_edge_keys(const int32_t* eu, const int32_t* ev, int64_t m, uint64_t* keys,
                      hipStream_t s);
void launch_edges_out(const uint64_t* keys, const uint32_t* idx, const float* ed2, int64_t m,
                      int32_t* u, int32_t* v, float* d2, hipStream_t s);
// Reference cellular matrix (findCell + CSR).
// geo: int32[16] = ext_dual[3], PC[3], pc[3], R, dim (host-computed, see emst_api.hip).
// dim = grid dimension, pstride = floats per point (3 for the 2.5D mode's 2D grid).
void launch_cell_grid(const float* resc, int64_t n, int dim, int pstride, const int32_t* geo, int32_t* cell_of,
                      uint64_t* keys, uint32_t* cell_count, hipStream_t s);
// learned seed radius (emst_mlp.hip)
// X: n x 8 model features, LS: n x log(leaf spacing s0)
void launch_radius_features(int dim, const float4* pts, const Node* nodes, int64_t L, LeafTable lt, int64_t n,
                            float* X, float* LS, hipStream_t s);
void launch_radius_targets(int K, int64_t n, const uint64_t* cand_key, const float* LS, float* T, hipStream_t s);
void launch_radius_seed(int64_t n, const float* model, const float* X, const float* LS, float* Y, float* R,
                        hipStream_t s);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                        hipStream_t s);

}  // namespace emst
