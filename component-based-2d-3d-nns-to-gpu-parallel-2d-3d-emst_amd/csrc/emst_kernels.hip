// emst_kernels.hip -- gfx950 kernels of the EMST engine.
//
// Stages (reference rows they implement, paths relative to the reference root):
//   bbox / params / rescale ..... InputRW.h:541-559, SolutionEMST.cu:305-379
//   Morton keys + radix sort .... the cell-grid build (count, wave-level scan,
//                                 scatter); cells are Morton prefixes
//   gather / tree ............... leaf cells of <= bucket points with tight boxes
//   cone search ................. NodeComputeOctantMD::computeOctants
//                                 (NodeEMST.h:1018-1296): per point, the nearest
//                                 point of each of the 2D x 2^(D-1) cones
//   round kernels ............... search (NodeEMST.h:1303-1384), component list +
//                                 min pair (NodeEMST.h:817-955), hooking
//                                 (EMSTOperators.h:1267-1303), flatten
//                                 (EMSTOperators.h:412-433)
//   weight ...................... K_EMST_evaluate_ST (EMSTOperators.h:262-292)
//   cell grid ................... ViewGridQuadMD::findCell (ViewGrid.h:339-359)
//
// Squared distances are fl(fl(dx*dx) + dy*dy) [+ dz*dz] with each "+ a*a"
// fused (__builtin_fmaf); the file is built with -ffp-contract=off so no other
// operation is contracted.
#include "emst_internal.h"

namespace emst {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float d2_of(float dx, float dy) { return __builtin_fmaf(dy, dy, dx * dx); }
__device__ __forceinline__ float d2_of(float dx, float dy, float dz)
{
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Order-preserving float <-> uint maps for atomic min/max of floats.
__device__ __forceinline__ uint32_t f2ord(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// maximum over each aligned group of 8 lanes, by DPP moves within rows (no
// LDS round trip): swap neighbours, swap pairs, then mirror the half row
// (lane i of the 8 reads lane 7 - i)
__device__ __forceinline__ uint32_t group8_max_u32(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));     // quad_perm [1,0,3,2]
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));     // quad_perm [2,3,0,1]
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false));    // row_half_mirror
    return v;
}

// ------------------------------------------------------------ clouds
// largest c with cl[c].pos <= p (the cloud holding global point p)
__device__ __forceinline__ int find_cloud(const Cloud* cl, int ncl, int64_t p)
{
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cl[mid].pos <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// largest c with cl[c].lbase <= g (the cloud holding global leaf g)
// This file is built with -fno-honor-nans (the kernels then skip the
// canonicalising moves the compiler puts before fminf / fmaxf inputs): every
// coordinate is checked finite before any search runs, and the tests for NaN
// and infinity look at the bits, passed through an empty asm statement so
// that the compiler cannot turn them back into floating-point class tests
// (which it would then fold under the no-NaN assumption).
__device__ __forceinline__ uint32_t opaque_bits(float x)
{
    uint32_t b = __float_as_uint(x);
    asm("" : "+v"(b));
    return b;
}
__device__ __forceinline__ bool finite_bits(float x)
{
    return (opaque_bits(x) & 0x7f800000u) != 0x7f800000u;
}
// x > 0 and finite
__device__ __forceinline__ bool positive_finite(float x)
{
    return opaque_bits(x) - 1u < 0x7f7fffffu;
}

__device__ __forceinline__ int find_cloud_leaf(const Cloud* cl, int ncl, int64_t g)
{
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cl[mid].lbase <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// sorted positions [s0, s1) of leaf slot j of a cloud (empty past its leaves)
__device__ __forceinline__ void leaf_range(const LeafTable& lt, int64_t lfirst, int64_t nl, int64_t cend, int64_t j,
                                           int64_t& s0, int64_t& s1)
{
    if (j < nl) {
        s0 = lt.start[lfirst + j];
        s1 = lt.start[lfirst + j + 1];
    } else {
        s0 = s1 = cend;
    }
}

// ------------------------------------------------------------ bbox (A1)
// Params.pmin/pmax hold order-mapped uints during the reduction.
__global__ void k_bbox_init(Params* P, int ncl)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    uint32_t* lo = reinterpret_cast<uint32_t*>(P[c].pmin);
    uint32_t* hi = reinterpret_cast<uint32_t*>(P[c].pmax);
    for (int k = 0; k < 3; ++k) { lo[k] = 0xffffffffu; hi[k] = 0u; }
    P[c].nonfinite = 0;
}

// Each wave walks a contiguous chunk of global points with per-lane
// accumulators for the chunk's current cloud, merged into P[cloud] (wave
// reduction, then order-mapped atomics) whenever the cloud changes.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bbox(const float* __restrict__ xyz, const Cloud* __restrict__ cl,
                                                int ncl, int64_t N, Params* __restrict__ P)
{
    const int lane = __lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / 64;
    const int64_t chunk = ((N + nwaves - 1) / nwaves + 63) / 64 * 64;
    const int64_t a = wave * chunk, b = min(N, a + chunk);
    uint32_t lo[D], hi[D];
    for (int k = 0; k < D; ++k) { lo[k] = 0xffffffffu; hi[k] = 0u; }
    uint32_t bad = 0;
    int cur = -1;
    auto flush = [&]() {
        for (int k = 0; k < D; ++k)
            for (int off = 32; off > 0; off >>= 1) {
                lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], off));
                hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], off));
            }
        const bool any_bad = __any(bad);
        if (lane == 0) {
            // the extremes only move outwards: a value already at least as far
            // out makes the atomic redundant (so for most waves once a few have run)
            uint32_t* plo = reinterpret_cast<uint32_t*>(P[cur].pmin);
            uint32_t* phi = reinterpret_cast<uint32_t*>(P[cur].pmax);
            for (int k = 0; k < D; ++k) {
                if (__atomic_load_n(&plo[k], __ATOMIC_RELAXED) > lo[k]) atomicMin(&plo[k], lo[k]);
                if (__atomic_load_n(&phi[k], __ATOMIC_RELAXED) < hi[k]) atomicMax(&phi[k], hi[k]);
            }
            if (any_bad) atomicOr(&P[cur].nonfinite, 1u);
        }
        for (int k = 0; k < D; ++k) { lo[k] = 0xffffffffu; hi[k] = 0u; }
        bad = 0;
    };
    for (int64_t i0 = a; i0 < b; i0 += 64) {
        const int64_t i = i0 + lane;
        const int c = i < b ? find_cloud(cl, ncl, i) : -1;
        uint64_t todo = __ballot(c >= 0);
        while (todo) {
            const int cc = __builtin_amdgcn_readfirstlane(__shfl(c, (int)__builtin_ctzll(todo)));
            if (cc != cur) {
                if (cur >= 0) flush();
                cur = cc;
            }
            const bool mine = c == cc;
            todo &= ~__ballot(mine);
            if (mine) {
                const float* q = xyz + i * D;
                for (int k = 0; k < D; ++k) {
                    const float v = q[k];
                    bad |= !finite_bits(v);
                    const uint32_t o = f2ord(v);
                    lo[k] = min(lo[k], o);
                    hi[k] = max(hi[k], o);
                }
            }
        }
    }
    if (cur >= 0) flush();
}

// ---------------------------------------------------------- params (A2)
// wg = (double)(pMax.x - pMin.x) etc. are float subtractions widened to double,
// _wnd / unitX in double (SolutionEMST.cu:305-342).  The rescaled extremes are
// the rescaled box corners: the map p -> (p - pMin) * unitXf is monotone.
// One thread per cloud; Morton quantisation uses mbits bits per axis.
__global__ void k_params(Params* Pa, const Cloud* __restrict__ cl, int ncl, int dim, int mbits)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    Params* P = Pa + c;
    uint32_t* plo = reinterpret_cast<uint32_t*>(P->pmin);
    uint32_t* phi = reinterpret_cast<uint32_t*>(P->pmax);
    const int64_t n = cl[c].n;
    for (int k = 0; k < 3; ++k) {
        if (k < dim && n > 0) {
            P->pmin[k] = ord2f(plo[k]);
            P->pmax[k] = ord2f(phi[k]);
        } else {
            P->pmin[k] = 0.f;
            P->pmax[k] = 0.f;
        }
    }
    double wg = (double)(P->pmax[0] - P->pmin[0]);
    double hg = (double)(P->pmax[1] - P->pmin[1]);
    double dg = dim >= 3 ? (double)(P->pmax[2] - P->pmin[2]) : 0.0;
    if (wg == 0) wg = 1;
    if (hg == 0) hg = 1;
    const double nn = (double)n;
    double wnd = sqrt(nn * wg / hg) * 2;
    if (dim >= 3 && dg != 0) wnd = pow((nn * wg * wg) / (hg * dg), 0.333);
    const double unit = (ceil(wnd) + 1) / wg;
    P->unit_x = unit;
    const float fu = (float)unit;
    P->unit_xf = fu;
    float mx = 0.f;
    for (int k = 0; k < 3; ++k) {
        P->newmax[k] = k < dim ? (P->pmax[k] - P->pmin[k]) * fu : 0.f;
        mx = fmaxf(mx, P->newmax[k]);
    }
    const double maxq = (double)((1ull << mbits) - 1);
    P->qscale = mx > 0 ? (float)(maxq / (double)mx) : 0.f;
}

// ------------------------------------------ rescale + Morton keys (A2, A5)
__device__ __forceinline__ uint64_t spread3(uint64_t x)
{
    x &= 0x1fffff;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint64_t spread2(uint64_t x)
{
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_rescale_morton(const float* __restrict__ xyz,
                                                           const Cloud* __restrict__ cl, int ncl,
                                                           int64_t N, int mbits,
                                                           const Params* __restrict__ Pa,
                                                           float* __restrict__ resc,
                                                           uint64_t* __restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = find_cloud(cl, ncl, i);
    const Params* P = Pa + c;
    const float* q = xyz + i * D;
    const float fu = P->unit_xf;
    const float qs = P->qscale;
    const uint32_t maxq = (uint32_t)((1ull << mbits) - 1);
    uint64_t code = 0;
    for (int k = 0; k < D; ++k) {
        float d = q[k] - P->pmin[k];                // PointE::operator- (Node.h:574-580)
        float v = d * fu;                           // PointE::operator*(float) (Node.h:588-593)
        resc[i * D + k] = v;
        float qf = v * qs;
        uint32_t qq = qf >= (float)maxq ? maxq : (uint32_t)qf;
        code |= (D == 3 ? spread3(qq) : spread2(qq)) << k;
    }
    if (ncl > 1) code |= (uint64_t)c << (mbits * D);
    keys[i] = code;
}

// the first active lane's v in every lane.  A load of one address by the
// whole wave is served in parts, so while other waves change the word its
// lanes can read different values; a value that steers the wave's control
// flow must then be one lane's (a partly active wave would also read zeros
// from the idle lanes in the shuffles of wave_min_u64).
__device__ __forceinline__ uint64_t wave_first_u64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// ------------------------------------------------------------------ scan
// Exclusive scan of uint32 in three passes: per-block scan of 1024 elements,
// scan of block totals (one block, chunked), add-back.
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = __lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total)
{
    const int lane = __lane_id(), w = threadIdx.x / 64, nw = blockDim.x / 64;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < nw; ++i) { uint32_t t = sh[i]; sh[i] = run; run += t; }
        sh[nw] = run;
    }
    __syncthreads();
    uint32_t r = sh[w] + inc - v;
    total = sh[nw];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void k_scan_local(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out, int64_t n,
                                                      uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[16];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, sh, total);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* sums, int64_t nb)
{
    __shared__ uint32_t sh[17];
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t total;
        uint32_t r = block_excl_scan(v, sh, total);
        if (i < nb) sums[i] = r + carry;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* __restrict__ out, int64_t n,
                                                    const uint32_t* __restrict__ sums)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const uint32_t add = sums[blockIdx.x];
    for (int j = threadIdx.x; j < kScanTile; j += kBlock)
        if (base + j < n) out[base + j] += add;
}

// ------------------------------------------------------------ radix sort
// 8-bit digits, tiles of 2048 keys per 256-thread block.  count: per-tile
// digit histogram (LDS atomics); scan over (digit, tile); scatter: each wave
// owns 512 consecutive keys of the tile and ranks equal digits with 8
// ballots, so the sort is stable; the tile is first sorted by digit in LDS
// and then written out in that order, so that the stores of consecutive
// threads fall on consecutive addresses of each digit's run.
constexpr int kRsTile = 2048;
constexpr int kRsWaveKeys = kRsTile / 4;
constexpr int kRsChunks = kRsWaveKeys / 64;

__global__ __launch_bounds__(kBlock) void k_rs_count(const uint64_t* __restrict__ keys, int64_t n,
                                                    int shift, uint32_t* __restrict__ hist,
                                                    int64_t nblk)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsTile / kBlock; ++j) {
        const int64_t i = base + j * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint64_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, int64_t n,
                                                      int shift, const uint32_t* __restrict__ hist,
                                                      int64_t nblk, int identity)
{
    __shared__ uint32_t cnt[4][256];        // per wave and digit; then the wave's next tile slot
    __shared__ uint32_t lstart[256];        // tile slot of each digit's first key
    __shared__ uint32_t gstart[256];        // output position of the tile's first key of each digit
    __shared__ uint32_t wsum[4];
    __shared__ uint64_t skey[kRsTile];
    __shared__ uint32_t sval[kRsTile];
    const int w = threadIdx.x / 64;
    const int lane = __lane_id();
    for (int k = threadIdx.x; k < 4 * 256; k += kBlock) (&cnt[0][0])[k] = 0;
    __syncthreads();
    const int64_t tbase = (int64_t)blockIdx.x * kRsTile;
    const int64_t wbase = tbase + (int64_t)w * kRsWaveKeys;
    uint64_t key[kRsChunks];
    uint32_t val[kRsChunks];
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        const int64_t i = wbase + c * 64 + lane;
        const bool valid = i < n;
        key[c] = valid ? kin[i] : 0ull;
        val[c] = identity ? (uint32_t)i : (valid ? vin[i] : 0u);
        if (valid) atomicAdd(&cnt[w][(key[c] >> shift) & 255u], 1u);
    }
    __syncthreads();
    {
        // exclusive scan of the tile's digit totals: within each wave, then
        // across the 4 waves' sums
        const int d = threadIdx.x;
        const uint32_t c0 = cnt[0][d], c1 = cnt[1][d], c2 = cnt[2][d], c3 = cnt[3][d];
        const uint32_t tot = c0 + c1 + c2 + c3;
        uint32_t inc = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)inc, off);
            if (lane >= off) inc += o;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t before = 0;
        for (int ww = 0; ww < w; ++ww) before += wsum[ww];
        const uint32_t ls = before + inc - tot;
        lstart[d] = ls;
        gstart[d] = hist[(int64_t)d * nblk + blockIdx.x];
        cnt[0][d] = ls;
        cnt[1][d] = ls + c0;
        cnt[2][d] = ls + c0 + c1;
        cnt[3][d] = ls + c0 + c1 + c2;
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        const int64_t i = wbase + c * 64 + lane;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(key[c] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(valid && bit);
            peers &= bit ? bb : ~bb;
        }
        if (valid) {
            const uint32_t slot = cnt[w][d] + (uint32_t)__popcll(peers & lt);
            skey[slot] = key[c];
            sval[slot] = val[c];
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0ull) cnt[w][d] += (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    const int tn = (int)min<int64_t>(kRsTile, n - tbase);
    for (int s = threadIdx.x; s < tn; s += kBlock) {
        const uint64_t kk = skey[s];
        const uint32_t d = (uint32_t)(kk >> shift) & 255u;
        const uint32_t pos = gstart[d] + (uint32_t)s - lstart[d];
        kout[pos] = kk;
        vout[pos] = sval[s];
    }
}

// -------------------------------------------------- gather sorted points
template <int D>
__global__ __launch_bounds__(kBlock) void k_gather_sorted(const float* __restrict__ resc,
                                                         const uint32_t* __restrict__ perm,
                                                         int64_t n, float4* __restrict__ pts,
                                                         int32_t* __restrict__ inv, uint32_t* __restrict__ ids)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    ids[s] = o;
    float4 p;
    p.x = resc[(int64_t)o * D + 0];
    p.y = resc[(int64_t)o * D + 1];
    p.z = D == 3 ? resc[(int64_t)o * D + 2] : 0.f;
    p.w = __uint_as_float(o);
    pts[s] = p;
    inv[o] = (int32_t)s;
}

// ------------------------------------------------ coincident-point groups
// Identical points share a Morton key, so each group lies inside one run of
// equal keys; the sort is stable over input order, so within a run positions
// ascend with input id and a group's first position holds its smallest id.
constexpr int kMixedRunMax = 1024;   // longest run holding >= 2 locations that is split
constexpr int kBigRun = 4096;        // longer runs are filled by the whole grid (k_dup_big)

__device__ __forceinline__ bool same_xyz(const float4& a, const float4& b)
{
    return a.x == b.x && a.y == b.y && a.z == b.z;
}

// append the lanes holding `take` to list, one atomic per wave
__device__ __forceinline__ void wave_append_u64(bool take, uint64_t v, uint64_t* list, uint32_t* counter)
{
    const uint64_t m = __ballot(take);
    if (m == 0ull) return;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (take) list[base + (uint32_t)__popcll(m & lanemask_lt())] = v;
}

// wave_append_u64 into a list of cap entries: one atomic per wave reserves
// the slots, and the lanes whose slot lies past the end are left out (the
// count may pass cap; its readers take min(count, cap), and every slot below
// cap is written by the lane that reserved it).  Returns this lane's taking
// part.  (A compare-and-swap loop that reserved only when all fitted queued
// the thousands of waves of one round on the count: 48 ms a round.)
__device__ __forceinline__ bool wave_append_capped_u64(bool take, uint64_t v, uint64_t* list, uint32_t* counter,
                                                       uint32_t cap)
{
    const uint64_t m = __ballot(take);
    if (m == 0ull) return false;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    const uint32_t slot = base + (uint32_t)__popcll(m & lanemask_lt());
    const bool in = take && slot < cap;
    if (in) list[slot] = v;
    return in;
}

// append the lanes holding `take` to list, one atomic per wave
__device__ __forceinline__ void wave_append(bool take, int32_t v, int32_t* list, uint32_t* counter)
{
    const uint64_t m = __ballot(take);
    if (m == 0ull) return;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (take) list[base + (uint32_t)__popcll(m & lanemask_lt())] = v;
}


// ------------------------------------------------------------ leaves
// Position i starts a leaf when it starts its cloud, or when the smallest
// prefix cell of the sorted keys that holds both i - 1 and i (the cell above
// their highest differing bit) has more than `bucket` points: its two halves
// then lie in different leaves.  The cell is contiguous around i, so two
// binary searches within `bucket` positions of i decide it.  Inside a run of
// equal keys longer than `bucket` a leaf starts every `bucket` points from
// the run's start.
__global__ __launch_bounds__(kBlock) void k_leaf_marks(const uint64_t* __restrict__ keys, int64_t n,
                                                      const Cloud* __restrict__ cl, int ncl, int bucket,
                                                      uint32_t* __restrict__ mark)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = find_cloud(cl, ncl, i);
    const int64_t c0 = cl[c].pos, c1 = cl[c].pos + cl[c].n;
    uint32_t m = 1u;
    if (i > c0) {
        const uint64_t a = keys[i - 1], b = keys[i];
        if (a != b) {
            const int h = 63 - __clzll(a ^ b);
            const uint64_t mask = h >= 63 ? 0ull : ~((2ull << h) - 1ull);
            const uint64_t pre = b & mask;
            int64_t lo = max(c0, i - bucket), hi = i - 1;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((keys[mid] & mask) == pre) hi = mid;
                else lo = mid + 1;
            }
            int64_t lo2 = i, hi2 = min(c1, i + bucket) - 1;
            while (lo2 < hi2) {
                const int64_t mid = (lo2 + hi2 + 1) >> 1;
                if ((keys[mid] & mask) == pre) lo2 = mid;
                else hi2 = mid - 1;
            }
            m = lo2 - lo + 1 > bucket ? 1u : 0u;
        } else {
            int64_t lo = i - 1, step = 1;
            while (lo - step >= c0 && keys[lo - step] == b) {
                lo -= step;
                step <<= 1;
            }
            int64_t a0 = max(c0, lo - step), b0 = lo;
            while (a0 < b0) {
                const int64_t mid = (a0 + b0) >> 1;
                if (keys[mid] == b) b0 = mid;
                else a0 = mid + 1;
            }
            const bool long_run = a0 + bucket < c1 && keys[a0 + bucket] == b;
            m = long_run && (i - a0) % bucket == 0 ? 1u : 0u;
        }
    }
    mark[i] = m;
}

// idx: exclusive scan of mark.  start[leaf] = its first position, of[p] =
// the leaf of p, start[leaves] = n.
__global__ __launch_bounds__(kBlock) void k_leaf_fill(const uint32_t* __restrict__ mark,
                                                     const uint32_t* __restrict__ idx, int64_t n,
                                                     int32_t* __restrict__ start, int32_t* __restrict__ of)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = idx[i] + mark[i] - 1u;
    of[i] = (int32_t)g;
    if (mark[i]) start[idx[i]] = (int32_t)i;
    if (i == n - 1) start[g + 1] = (int32_t)n;
}

// First leaf of every cloud (every non-empty cloud starts a leaf), and the
// count.  layout: also lay the clouds' hierarchies out in the cloud table as
// the host does after reading the counts (L = the power of two at or above
// the cloud's leaf count, lbase = the sum of the earlier clouds' L), inside
// slot counts the host fixed beforehand (the table's L on entry);
// C->invalid flags a cloud with more leaves than those slots, or with a NaN
// or Inf coordinate.  One block.
__global__ __launch_bounds__(1024) void k_cloud_leaves(Cloud* __restrict__ cl, int ncl,
                                                      const uint32_t* __restrict__ mark,
                                                      const uint32_t* __restrict__ idx, int64_t n,
                                                      int32_t* __restrict__ out, const Params* __restrict__ P,
                                                      Counters* C, int layout)
{
    __shared__ int64_t part[1024];
    if (threadIdx.x == 0) C->invalid = 0u;
    const int32_t total = n > 0 ? (int32_t)(idx[n - 1] + mark[n - 1]) : 0;
    for (int c = threadIdx.x; c <= ncl; c += blockDim.x)
        out[c] = (c < ncl && cl[c].pos < n) ? (int32_t)idx[cl[c].pos] : total;
    __syncthreads();
    if (!layout) return;
    // each thread a run of consecutive clouds
    const int per = (ncl + (int)blockDim.x - 1) / (int)blockDim.x;
    const int c0 = min(ncl, (int)threadIdx.x * per), c1 = min(ncl, c0 + per);
    int64_t sum = 0;
    for (int c = c0; c < c1; ++c) {
        const int64_t nl = (int64_t)out[c + 1] - out[c];
        if (nl > cl[c].L || P[c].nonfinite) atomicOr(&C->invalid, 1u);
        int64_t L = 1;
        while (L < nl) L <<= 1;
        cl[c].lfirst = out[c];
        cl[c].nl = nl;
        cl[c].L = L;
        sum += L;
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const int64_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for (int c = c0; c < c1; ++c) {
        cl[c].lbase = run;
        run += cl[c].L;
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// rep[] = -1 everywhere; the first position of every run of >= 2 equal keys is queued
__global__ __launch_bounds__(kBlock) void k_dup_runs(const uint64_t* __restrict__ keys, int64_t n,
                                                    int32_t* __restrict__ rep, int32_t* __restrict__ runs,
                                                    Counters* __restrict__ C)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) C->dup_big = 0;
    if (i >= n) return;
    rep[i] = -1;
    const uint64_t k = keys[i];
    const bool head = i == 0 || keys[i - 1] != k;
    if (head && i + 1 < n && keys[i + 1] == k) runs[atomicAdd(&C->dup_runs, 1u)] = (int32_t)i;
}

// One wave per run (grid-stride over the queued runs).  A run at one location
// is one group; a run over several locations (distinct points closer than a
// Morton cell) of at most kMixedRunMax points is split by pairwise
// comparison; a longer one sets dup_overflow, which turns grouping off.
// A run of more than kBigRun points (a heavily repeated location: one wave
// would walk it alone while the grid idles) gets its group slot and list
// range here and is checked and filled by k_dup_big; its (start, end, list
// base) go to the back of `runs`, entry k at runs[n - 3k - 3 ..]: the front
// holds at most n / 2 run heads and the back at most 3n / (kBigRun + 1).
__global__ __launch_bounds__(kBlock) void k_dup_group(const uint64_t* __restrict__ keys,
                                                     const float4* __restrict__ pts, int64_t n,
                                                     int32_t* __restrict__ rep,
                                                     int32_t* __restrict__ runs,
                                                     int32_t* __restrict__ list, int32_t* __restrict__ gslot,
                                                     Counters* __restrict__ C)
{
    const int lane = __lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
    const int64_t nruns = C->dup_runs;
    for (int64_t w = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64; w < nruns; w += nwaves) {
        const int64_t s = runs[w];
        const uint64_t key = keys[s];
        // run end: the first larger key (the keys are sorted)
        int64_t lo = s + 1, hi = n;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (keys[mid] == key) lo = mid + 1;
            else hi = mid;
        }
        const int64_t e = lo;
        if (e - s > kBigRun) {
            if (lane == 0) {
                gslot[s] = (int32_t)atomicAdd(&C->dup_groups, 1u);
                const uint32_t base = atomicAdd(&C->dup_points, (uint32_t)(e - s));
                const int64_t k = (int64_t)atomicAdd(&C->dup_big, 1u);
                runs[n - 3 * k - 3] = (int32_t)s;
                runs[n - 3 * k - 2] = (int32_t)e;
                runs[n - 3 * k - 1] = (int32_t)base;
            }
            continue;
        }
        const float4 head = pts[s];
        bool mixed = false;
        // four independent loads per lane per trip
        for (int64_t b = s; b < e; b += 256) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t i = b + 64 * u + lane;
                if (i < e) mixed |= !same_xyz(pts[i], head);
            }
        }
        if (__ballot(mixed) == 0ull) {
            // one group: its list entries are reserved in one step
            uint32_t base = 0;
            if (lane == 0) {
                gslot[s] = (int32_t)atomicAdd(&C->dup_groups, 1u);
                base = atomicAdd(&C->dup_points, (uint32_t)(e - s));
            }
            base = (uint32_t)__shfl((int)base, 0);
            for (int64_t i = s + lane; i < e; i += 64) {
                rep[i] = (int32_t)s;
                list[base + (uint32_t)(i - s)] = (int32_t)i;
            }
        } else if (e - s <= kMixedRunMax) {
            for (int64_t b = s; b < e; b += 64) {
                const int64_t i = b + lane;
                const float4 me = i < e ? pts[i] : head;
                int64_t first = -1;
                int twins = 0;
                for (int64_t j = s; j < e; ++j) {
                    if (same_xyz(pts[j], me)) {
                        if (first < 0) first = j;
                        ++twins;
                    }
                }
                const bool grp = i < e && twins >= 2;
                if (grp) rep[i] = (int32_t)first;
                if (grp && first == i) gslot[i] = (int32_t)atomicAdd(&C->dup_groups, 1u);
                wave_append(grp, (int32_t)i, list, &C->dup_points);
            }
        } else if (lane == 0) {
            atomicOr(&C->dup_overflow, 1u);
        }
    }
}

// The long runs queued by k_dup_group, each spread over the whole grid: one
// group when every point sits at the run head's location, else a mixed run
// too long to split (grouping off).
__global__ __launch_bounds__(kBlock) void k_dup_big(const float4* __restrict__ pts, int64_t n,
                                                   int32_t* __restrict__ rep, const int32_t* __restrict__ runs,
                                                   int32_t* __restrict__ list, Counters* __restrict__ C)
{
    const uint32_t nb = C->dup_big;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (uint32_t k = 0; k < nb; ++k) {
        const int64_t s = runs[n - 3 * (int64_t)k - 3], e = runs[n - 3 * (int64_t)k - 2];
        const uint32_t base = (uint32_t)runs[n - 3 * (int64_t)k - 1];
        const float4 head = pts[s];
        bool mixed = false;
        for (int64_t i = s + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
            mixed |= !same_xyz(pts[i], head);
            rep[i] = (int32_t)s;
            list[base + (uint32_t)(i - s)] = (int32_t)i;
        }
        if (__ballot(mixed) != 0ull && __lane_id() == 0) atomicOr(&C->dup_overflow, 1u);
    }
}

__global__ void k_dup_finish(Counters* C)
{
    C->grouped = (C->dup_points > 0u && C->dup_overflow == 0u) ? 1u : 0u;
}

// ---------------------------------------------------------------- tree
// A leaf box covers the leaf's candidates: while coincident-point grouping is
// on, group members are never candidates (their representative, the group's
// smallest id, wins every tie at their common location), so they are left out
// and a leaf of members only keeps the empty box (+inf, -inf), which no node
// test accepts.
__global__ __launch_bounds__(kBlock) void k_leaves(const float4* __restrict__ pts,
                                                  const Cloud* __restrict__ cl, int ncl, int64_t nleaves,
                                                  LeafTable lt, Node* __restrict__ nodes,
                                                  const int32_t* __restrict__ rep)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nleaves) return;
    const int c = find_cloud_leaf(cl, ncl, g);
    const int64_t j = g - cl[c].lbase;          // the cloud's leaf slot j
    if (j >= cl[c].L) return;               // past the cloud's slots (one-launch capacity)
    const bool grouped = rep != nullptr;
    Node nd;
    for (int k = 0; k < 3; ++k) { nd.lo[k] = INFINITY; nd.hi[k] = -INFINITY; }
    nd.pad[0] = nd.pad[1] = 0.f;
    int64_t s0, s1;
    leaf_range(lt, cl[c].lfirst, cl[c].nl, cl[c].pos + cl[c].n, j, s0, s1);
    for (int64_t s = s0; s < s1; ++s) {
        if (grouped) {
            const int32_t r = rep[s];
            if (r >= 0 && r != (int32_t)s) continue;
        }
        const float4 p = pts[s];
        nd.lo[0] = fminf(nd.lo[0], p.x); nd.hi[0] = fmaxf(nd.hi[0], p.x);
        nd.lo[1] = fminf(nd.lo[1], p.y); nd.hi[1] = fmaxf(nd.hi[1], p.y);
        nd.lo[2] = fminf(nd.lo[2], p.z); nd.hi[2] = fmaxf(nd.hi[2], p.z);
    }
    nodes[2 * cl[c].lbase + cl[c].L + j] = nd;
}

// heap level [lo, 2 lo) of every cloud with more than lo leaves
__global__ __launch_bounds__(kBlock) void k_level(Node* __restrict__ nodes, const Cloud* __restrict__ cl,
                                                 int ncl, int64_t lo)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ncl * lo) return;
    const int c = (int)(t / lo);
    if (lo >= cl[c].L) return;
    Node* base = nodes + 2 * cl[c].lbase;
    const int64_t i = lo + t % lo;
    const Node a = base[2 * i], b = base[2 * i + 1];
    Node nd;
    for (int k = 0; k < 3; ++k) {
        nd.lo[k] = fminf(a.lo[k], b.lo[k]);
        nd.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    nd.pad[0] = nd.pad[1] = 0.f;
    base[i] = nd;
}

// --------------------------------------------------------- cone search
// One wave = 64 query points (consecutive in Morton order, or a query list).
// The wave walks the binary hierarchy together (packet traversal): a node is
// entered when any lane may still improve a cone from it; leaf points are
// read with wave-uniform (scalar) loads and every lane tests them against its
// own query.  Per lane, the best key (d2 bits << 32 | input id) of each cone
// sits in LDS; the key order is (d2, id), which for a fixed query equals
// EMST_isInf's (d2, min id, max id) order (macros_cuda.h:37).
//
// Cone membership follows NodeEMST.h:1204-1233 on the float difference
// q - p: face (axis, sign) for every axis whose |d| equals the max (sign from
// d >= 0, so -0 is '+'), and every slab of that face whose non-axis signs
// agree, where a zero coordinate is in both halves.
//
// Node test: for each face, the smallest float d2 any point of the box can
// have inside the face's pyramid (|d_axis| >= |d_other|), computed with the
// same rounded operations on the box corners, so it never exceeds a point's
// own d2; the face is kept while that bound is <= the face's worst cone best
// (empty cone = +inf).  Equality is kept because a tie can still win on id.
template <int D>
struct ConeCfg {
    static constexpr int NF = 2 * D;
    static constexpr int NS = 1 << (D - 1);
    static constexpr int K = NF * NS;
};

// allowed halves of one non-axis coordinate: bit0 = '+' half, bit1 = '-' half
__device__ __forceinline__ uint32_t halves(float d)
{
    return d > 0.f ? 1u : (d < 0.f ? 2u : 3u);
}

template <int D>
__device__ __forceinline__ uint32_t cone_mask(float dx, float dy, float dz)
{
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const uint32_t hx = halves(dx), hy = halves(dy);
    uint32_t mask = 0;
    if (D == 2) {
        const float m = fmaxf(ax, ay);
        if (ax >= m) mask |= hy << (2 * (dx < 0.f ? 1 : 0));
        if (ay >= m) mask |= hx << (2 * (dy < 0.f ? 3 : 2));
    } else {
        const uint32_t hz = halves(dz);
        const float m = fmaxf(ax, fmaxf(ay, az));
        // slab s = bit0 (first non-axis), bit1 (second); half code b -> bit value
        // expand(h1, h2): slabs {s : (h1 >> (s&1)) & 1 && (h2 >> (s>>1)) & 1}
        auto expand = [](uint32_t h1, uint32_t h2) -> uint32_t {
            return ((h2 & 1u) ? h1 : 0u) | ((h2 & 2u) ? (h1 << 2) : 0u);
        };
        if (ax >= m) mask |= expand(hy, hz) << (4 * (dx < 0.f ? 1 : 0));
        if (ay >= m) mask |= expand(hx, hz) << (4 * (dy < 0.f ? 3 : 2));
        if (az >= m) mask |= expand(hx, hy) << (4 * (dz < 0.f ? 5 : 4));
    }
    return mask;
}

// Node test.  For every point q of the box, |d_k| >= m_k on each axis, m_k the
// smallest |d_k| over the box; a half-axis (d_k >= 0 or d_k <= 0) that the box
// reaches always has that same minimum (0 when the box straddles p's
// coordinate, else the one reachable half).  So inside face (axis a, sign) the
// box's points satisfy |d_a| >= max(smallest |d_a| on the face's side, m_o1,
// m_o2) and |d_oi| >= m_oi, one bound shared by all the face's slabs the box
// reaches.  The float d2 formula is monotone in each |d|, so the same rounded
// operations on these bounds never exceed any q's own d2.  The box is needed
// when, for some face, that bound is <= the largest bound among the face's
// reachable open slabs, and its farthest point lies beyond `inner` (points
// with d2 <= inner were examined in an earlier pass).  Bounds are int32 d2
// bit patterns: -1 = cone closed, INT_MAX = open without candidate.
//
// Faces are strict where the cone is: a '-' face holds d_a < 0 only, so the
// box must reach below the query on that axis; a '+' face holds d_a > 0 or
// the zero vector, and with strict_plus (coincident points never candidates)
// d_a > 0 only.
//
// D = distance dimension, DC = cone dimension.  DC = 2 with D = 3 (the 2.5D
// mode) classifies by the xy offset only, and every face bound also carries
// the box's smallest |dz|.
template <int D, int DC>
__device__ __forceinline__ bool node_needed(const Node& nd, float px, float py, float pz,
                                            const int32_t* bound, int32_t maxb, int32_t inner,
                                            bool strict_plus)
{
    float l[3], h[3], m[3], u[3];
    bool fp[3], fn[3];
    l[0] = nd.lo[0] - px; h[0] = nd.hi[0] - px;
    l[1] = nd.lo[1] - py; h[1] = nd.hi[1] - py;
    if (D == 3) { l[2] = nd.lo[2] - pz; h[2] = nd.hi[2] - pz; }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        m[k] = fmaxf(fmaxf(l[k], -h[k]), 0.f);
        u[k] = fmaxf(-l[k], h[k]);
        fp[k] = h[k] >= 0.f;
        fn[k] = l[k] <= 0.f;
    }
    {
        const float lb = D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]);
        if ((int32_t)__float_as_uint(lb) > maxb) return false;
        const float far = D == 3 ? d2_of(u[0], u[1], u[2]) : d2_of(u[0], u[1]);
        if ((int32_t)__float_as_uint(far) <= inner) return false;
    }
    bool need = false;
#pragma unroll
    for (int f = 0; f < 2 * DC; ++f) {
        const int a = f / 2;
        const bool neg = f & 1;
        const int o1 = a == 0 ? 1 : 0;
        const int o2 = a == 2 ? 1 : 2;
        const float o = DC == 3 ? fmaxf(m[o1], m[o2]) : m[o1];
        const float reach = neg ? -l[a] : h[a];
        const float am = fmaxf(neg ? -h[a] : l[a], o);
        float v[3];
        v[a] = am;
        v[o1] = m[o1];
        if (D == 3) v[o2] = m[o2];      // DC = 2: o2 is the z axis, outside the cone
        const float lb = D == 3 ? d2_of(v[0], v[1], v[2]) : d2_of(v[0], v[1]);
        const int32_t* b = bound + f * (1 << (DC - 1));
        int32_t bm;
        if (DC == 3) {
            bm = max(max((fp[o1] && fp[o2]) ? b[0] : -1, (fn[o1] && fp[o2]) ? b[1] : -1),
                     max((fp[o1] && fn[o2]) ? b[2] : -1, (fn[o1] && fn[o2]) ? b[3] : -1));
        } else {
            bm = max(fp[o1] ? b[0] : -1, fn[o1] ? b[1] : -1);
        }
        const bool open_side = (neg || strict_plus) ? reach > 0.f : true;
        need |= open_side && (reach >= o) && ((int32_t)__float_as_uint(lb) <= bm);
    }
    return need;
}

// best key -> int32 bound pattern (empty cone = INT_MAX)
__device__ __forceinline__ int32_t key_bound(unsigned long long k)
{
    const uint32_t hi = (uint32_t)(k >> 32);
    return hi == 0xffffffffu ? 0x7fffffff : (int32_t)hi;
}

constexpr int kMaxCappedPasses = 6;

// Cone search, exact, by iterative deepening (the tree form of the Elias ring
// expansion).  Pass k caps every open cone's bound at R_k^2 (R_0 = the seed,
// R_{k+1} = 2 R_k); after the pass a cone whose best d2 is <= R_k^2 is exact
// and closes; the pass after kMaxCappedPasses has no cap, so every cone ends
// exact.  The seed only decides how much work each pass does, never the result.
//
// Seeds: seed_r (per sorted point) when given, else seed_scale times the mean
// spacing of the point's own leaf cell ((box volume / points)^(1/D)).
//
// Passes [pass_lo, pass_hi] run here.  With open_out, a query still holding
// an open cone after pass_hi stores its cone bests so far (cand_key) and its
// lane bit in open_out[packet]; a later launch (pass_lo = pass_hi + 1, nq_dev
// = the number of such queries, counted on the device) reloads them and goes
// on, in one of two forms chosen on the device: when at most compact_max
// queries are open, packets of the compacted list qlist (after pass 0 most
// queries are done, so the later passes then run on full waves instead of
// dragging a few live lanes through whole traversals); otherwise each packet
// resumes its own open lanes (resume_mask).  Resuming after pass k rebuilds
// the state exactly: R_k by the same float products, and a cone is open after
// pass k iff its best bound exceeds R_k^2.
// Occupancy: the search waits on dependent box and point loads, so it pays;
// with 8 cones the kernel fits 5 waves per SIMD in 96 VGPRs without spills
// (+8% on the 2D power-law cloud, +9% in the 2.5D mode).  With 24 cones the
// cone bests take 12 KB of LDS per 64-query block, which allows 12 blocks per
// CU, 3 waves per SIMD, whatever the registers; the attribute asks for that
// (160 VGPRs, no VGPR spills).
template <int D, int DC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DC == 3 ? 3 : 5))) void k_cone_search(const float4* __restrict__ pts,
                                                   const Node* __restrict__ nodes,
                                                   const Cloud* __restrict__ clouds, int ncl,
                                                   int bucket, LeafTable lt, int64_t n,
                                                   const int32_t* __restrict__ qlist, int64_t nq,
                                                   const int32_t* __restrict__ label,
                                                   int check_labels,
                                                   const int32_t* __restrict__ inv,
                                                   const float* __restrict__ seed_r,
                                                   float seed_scale,
                                                   uint64_t* __restrict__ cand_key,
                                                   int32_t* __restrict__ cand_pos,
                                                   int32_t* __restrict__ revive,
                                                   int32_t* __restrict__ first_corr,
                                                   uint64_t* __restrict__ first_key,
                                                   Counters* __restrict__ ctr, int collect_stats,
                                                   DupState dup, int pass_lo, int pass_hi,
                                                   uint64_t* __restrict__ open_out,
                                                   const uint32_t* __restrict__ nq_dev,
                                                   const uint64_t* __restrict__ resume_mask,
                                                   int64_t compact_max)
{
    using C = ConeCfg<DC>;
    bool compacted = true;
    if (nq_dev) {
        nq = *nq_dev;
        compacted = nq <= compact_max;
        if (compacted ? (int64_t)blockIdx.x * 64 >= nq : resume_mask[blockIdx.x] == 0ull)
            return;     // nothing to resume in this packet
    }
    __shared__ unsigned long long best[C::K][64];
    __shared__ int32_t stk_id[64];
    __shared__ int32_t stk_ver[64];
    __shared__ unsigned long long stk_mask[64];
    const int lane = __lane_id();
    const int64_t qi = (int64_t)blockIdx.x * 64 + lane;
    bool active;
    int64_t p;
    if (compacted) {
        active = qi < nq;
        p = active ? (qlist ? (int64_t)qlist[qi] : qi) : 0;
    } else {
        active = (resume_mask[blockIdx.x] >> lane) & 1ull;     // set bits are points of this packet
        p = active ? qi : 0;
    }
    // grouping on: group members do not search (k_dup_fill gives them their
    // representative's cones) and no point takes a candidate at its own
    // location; otherwise only the point itself is skipped
    const bool grouped = dup.rep != nullptr;
    // strict '+' faces need the '+' cones to hold no point but the query's
    // own location; with xy wedges over 3D points they also hold the points
    // straight above and below it
    const bool strict_plus = grouped && DC == D;
    bool is_rep = false;
    if (grouped && active) {
        const int32_t r = dup.rep[p];
        active = r < 0 || r == (int32_t)p;
        is_rep = r == (int32_t)p;
    }
    bool underflow = false;
    const float4 me = pts[p];
    const float px = me.x, py = me.y, pz = me.z;
    const int32_t mylab = (check_labels && active) ? label[p] : 0;
#pragma unroll
    for (int c = 0; c < C::K; ++c) best[c][lane] = (pass_lo > 0 && active) ? cand_key[(int64_t)c * n + p] : kEmptyKey;

    uint32_t st_nodes = 0, st_leaves = 0, st_pts = 0, st_pairs = 0;
    // one traversal per cloud among the wave's queries (a packet of a single
    // cloud, the usual case, makes one)
    const int my_cl = active ? find_cloud(clouds, ncl, p) : -1;
    uint32_t open_final = 0;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int cl = __builtin_amdgcn_readfirstlane(__shfl(my_cl, (int)__builtin_ctzll(pending)));
        const bool mine = my_cl == cl;
        pending &= ~__ballot(mine);
        const int64_t cend = clouds[cl].pos + clouds[cl].n;
        const int64_t L = clouds[cl].L, lfirst = clouds[cl].lfirst, nl = clouds[cl].nl;
        const Node* __restrict__ tree = nodes + 2 * clouds[cl].lbase;

        // seed radius
        float r = 1.f;
        if (!mine) {
            // not this traversal's query
        } else if (seed_r) {
            r = seed_r[p];
            if (!positive_finite(r)) r = 1e-6f;
        } else {
            const int64_t lf = lt.of[p];
            const Node own = tree[L + (lf - lfirst)];
            const float cnt = (float)(lt.start[lf + 1] - lt.start[lf]);
            float vol = 1.f, ext = 0.f;
            for (int k = 0; k < D; ++k) {
                const float e = own.hi[k] - own.lo[k];
                vol *= e;
                ext = fmaxf(ext, e);
            }
            float sp = vol > 0.f ? powf(vol / cnt, 1.f / D) : ext / cnt;
            if (!positive_finite(sp)) sp = 1e-6f;
            r = seed_scale * sp;
        }
        float r2 = r * r;
        uint32_t open = mine ? (uint32_t)((1ull << C::K) - 1) : 0u;
        int32_t inner = -1;
        for (int k = 0; k < pass_lo; ++k) {
            inner = (int32_t)__float_as_uint(r2);
            r2 *= 4.f;
        }
        if (pass_lo > 0) {
            open = 0u;
#pragma unroll
            for (int c = 0; c < C::K; ++c)
                if (mine && key_bound(best[c][lane]) > inner) open |= 1u << c;
        }
        int32_t bound[C::K];
        int32_t maxb = -1;

        // packet centre (for near-first child order)
        float cx = mine ? px : 0.f, cy = mine ? py : 0.f, cz = mine ? pz : 0.f;
        float na = mine ? 1.f : 0.f;
        for (int off = 32; off > 0; off >>= 1) {
            cx += __shfl_xor(cx, off);
            cy += __shfl_xor(cy, off);
            cz += __shfl_xor(cz, off);
            na += __shfl_xor(na, off);
        }
        cx /= na; cy /= na; cz /= na;
        auto cdist = [&](const Node& x) {
            const float dxx = fmaxf(fmaxf(x.lo[0] - cx, cx - x.hi[0]), 0.f);
            const float dyy = fmaxf(fmaxf(x.lo[1] - cy, cy - x.hi[1]), 0.f);
            const float dzz = fmaxf(fmaxf(x.lo[2] - cz, cz - x.hi[2]), 0.f);
            return dxx * dxx + dyy * dyy + dzz * dzz;
        };

        for (int pass = pass_lo; pass <= pass_hi; ++pass) {
            if (__ballot(open != 0u) == 0ull) break;
            const bool last = pass == kMaxCappedPasses;
            const int32_t cap = last ? 0x7fffffff : (int32_t)__float_as_uint(r2);
            auto refresh = [&]() {
                maxb = -1;
#pragma unroll
                for (int c = 0; c < C::K; ++c) {
                    const int32_t b = ((open >> c) & 1u) ? min(key_bound(best[c][lane]), cap) : -1;
                    bound[c] = b;
                    maxb = max(maxb, b);
                }
            };
            refresh();
            int sp = 0;
            int ver = 0;        // bumps whenever a leaf tightened the bounds
            if (lane == 0) { stk_id[0] = 1; stk_ver[0] = -1; }
            sp = 1;
            __builtin_amdgcn_wave_barrier();
            while (sp > 0) {
                --sp;
                const int node = __builtin_amdgcn_readfirstlane(stk_id[sp]);
                const int pver = __builtin_amdgcn_readfirstlane(stk_ver[sp]);
                // a node tested against the current bounds when pushed keeps its
                // lane mask; after a leaf tightened the bounds it is tested again
                bool need;
                if (pver == ver) {
                    need = (stk_mask[sp] >> lane) & 1ull;
                } else {
                    const Node nd = tree[node];
                    need = open != 0u && node_needed<D, DC>(nd, px, py, pz, bound, maxb, inner, strict_plus);
                    ++st_nodes;
                }
                const uint64_t needm = __ballot(need);
                if (needm == 0ull) continue;
                if (node >= L) {
                    int64_t s0, s1;
                    leaf_range(lt, lfirst, nl, cend, node - L, s0, s1);
                    const int cnt = (int)(s1 - s0);
                    ++st_leaves;
                    st_pts += (uint32_t)cnt;
                    st_pairs += (uint32_t)cnt * (uint32_t)__popcll(needm);
                    // pair-parallel: lane = (query slot, leaf point); a chunk covers
                    // 64 / bucket needed queries x the leaf's points
                    const int j = lane & (bucket - 1);
                    const int qslot = lane / bucket;
                    const int qpc = 64 / bucket;
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (j < cnt) v = pts[s0 + j];
                    const int64_t s = s0 + j;
                    // one (query, leaf point) pair per lane: distance, cone, and
                    // an LDS atomic into the query's cone best
                    auto pair = [&](int myq, float qx, float qy, float qz, int32_t qmaxb, int32_t qinner,
                                    uint32_t qopen, int32_t qpos, int32_t qlab) {
                        const float dx = v.x - qx, dy = v.y - qy, dz = v.z - qz;
                        const float d2 = D == 3 ? d2_of(dx, dy, dz) : d2_of(dx, dy);
                        const int32_t db = (int32_t)__float_as_uint(d2);
                        const bool here = dx == 0.f && dy == 0.f && dz == 0.f;
                        bool take = myq >= 0 && j < cnt && (grouped ? !here : s != (int64_t)qpos) &&
                                    db <= qmaxb && db > qinner;
                        // distinct points at float distance 0 (underflow) break the
                        // grouping's premise; the caller then searches ungrouped
                        if (grouped && take && d2 == 0.f) underflow = true;
                        if (check_labels && take && d2 == 0.f) take = label[s] != qlab;
                        if (take) {
                            uint32_t mask = cone_mask<DC>(dx, dy, dz) & qopen;
                            const unsigned long long key =
                                ((unsigned long long)__float_as_uint(d2) << 32) | __float_as_uint(v.w);
                            while (mask) {
                                const int c = __builtin_ctz(mask);
                                mask &= mask - 1;
                                atomicMin(&best[c][myq], key);
                            }
                        }
                    };
                    uint64_t rem = needm;
                    if (qpc == 1) {
                        // every lane works on the same query: broadcast its state with
                        // v_readlane (a scalar register) instead of an LDS permute
                        while (rem) {
                            const int q = (int)__builtin_ctzll(rem);
                            rem &= rem - 1;
                            auto rl = [&](int x) { return __builtin_amdgcn_readlane(x, q); };
                            pair(q, __int_as_float(rl(__float_as_int(px))), __int_as_float(rl(__float_as_int(py))),
                                 __int_as_float(rl(__float_as_int(pz))), rl(maxb), rl(inner), (uint32_t)rl((int)open),
                                 rl((int32_t)p), check_labels ? rl(mylab) : 0);
                        }
                    } else {
                        while (rem) {
                            // the qslot-th remaining needed lane is this lane's query
                            int myq = -1;
                            for (int k = 0; k < qpc; ++k) {
                                const int l = rem ? (int)__builtin_ctzll(rem) : -1;
                                if (rem) rem &= rem - 1;
                                if (k == qslot) myq = l;
                            }
                            const int src = myq >= 0 ? myq : 0;
                            pair(myq, __shfl(px, src), __shfl(py, src), __shfl(pz, src), __shfl(maxb, src),
                                 __shfl(inner, src), (uint32_t)__shfl((int)open, src), __shfl((int32_t)p, src),
                                 check_labels ? __shfl(mylab, src) : 0);
                        }
                    }
                    if (need) refresh();
                    ++ver;
                } else {
                    // the next boxes down: the children when they are leaves, else
                    // the four grandchildren (a descent then waits on half as many
                    // dependent loads; the four loads are independent)
                    const bool leafkids = 2 * (int64_t)node >= L;
                    const int first = leafkids ? 2 * node : 4 * node;
                    const int nk = leafkids ? 2 : 4;
                    Node bx[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (i < nk) bx[i] = tree[first + i];
                    uint64_t mk[4];
                    float dk[4];
                    int id[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        mk[i] = i < nk ? __ballot(open != 0u && node_needed<D, DC>(bx[i], px, py, pz, bound,
                                                                                     maxb, inner, strict_plus))
                                       : 0ull;
                        dk[i] = mk[i] ? cdist(bx[i]) : -1.f;
                        id[i] = first + i;
                    }
                    st_nodes += (uint32_t)nk;
                    // farthest pushed first, so the nearest is popped next
                    auto ce = [&](int a, int b) {
                        if (dk[a] < dk[b]) {
                            const float td = dk[a]; dk[a] = dk[b]; dk[b] = td;
                            const uint64_t tm = mk[a]; mk[a] = mk[b]; mk[b] = tm;
                            const int ti = id[a]; id[a] = id[b]; id[b] = ti;
                        }
                    };
                    ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2);
                    if (lane == 0) {
                        int t = sp;
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (mk[i]) { stk_id[t] = id[i]; stk_ver[t] = ver; stk_mask[t] = mk[i]; ++t; }
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) sp += mk[i] != 0ull ? 1 : 0;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // close the cones whose best lies within this pass's radius
#pragma unroll
            for (int c = 0; c < C::K; ++c)
                if (((open >> c) & 1u) && key_bound(best[c][lane]) <= cap) open &= ~(1u << c);
            inner = cap;
            r2 *= 4.f;
        }
        if (mine) open_final = open;
    }
    if (collect_stats && lane == 0) {
        atomicAdd(&ctr->packets, 1ull);
        atomicAdd(&ctr->nodes, (unsigned long long)st_nodes);
        atomicAdd(&ctr->leaves, (unsigned long long)st_leaves);
        atomicAdd(&ctr->leaf_points, (unsigned long long)st_pts);
        atomicAdd(&ctr->pairs, (unsigned long long)st_pairs);
    }
    if (__ballot(underflow) != 0ull && lane == 0) atomicOr(&ctr->underflow, 1u);
    if (open_out) {
        const bool unfinished = active && open_final != 0u;
        const uint64_t om = __ballot(unfinished);
        if (lane == 0) open_out[blockIdx.x] = om;
        if (unfinished) {
            // the bests so far, for the launch that resumes this query
            for (int c = 0; c < C::K; ++c) cand_key[(int64_t)c * n + p] = best[c][lane];
            return;
        }
    }
    if (!active) return;
    if (revive && revive[p] == -2) revive[p] = -1;
    const int64_t gs = is_rep ? (int64_t)dup.gslot[p] : 0;
    // cone-major layout [K][n]: lanes write consecutive words
    uint64_t kmin = kEmptyKey;
    int32_t pmin = -1;
    for (int c = 0; c < C::K; ++c) {
        const uint64_t k = best[c][lane];
        const int32_t q = (k == kEmptyKey) ? -1 : inv[(uint32_t)k];
        cand_key[(int64_t)c * n + p] = k;
        cand_pos[(int64_t)c * n + p] = q;
        if (k < kmin) { kmin = k; pmin = q; }
        // a representative keeps its '+' cones for the rounds in which its
        // group has become one component (k_dup_fill)
        if (is_rep && ((c / C::NS) & 1) == 0) {
            const int64_t b = (int64_t)((c / C::NS / 2) * C::NS + c % C::NS) * dup.ngroups + gs;
            dup.bkey[b] = k;
            dup.bpos[b] = q;
        }
    }
    // first search (all points singletons): the round-1 target is the overall
    // cone minimum (NodeEMST.h:1338-1381 with every label distinct)
    if (first_corr) {
        first_corr[p] = pmin >= 0 ? pmin : -2;
        first_key[p] = kmin;
    }
}

// Queries left open by a split cone search, in position order: per-packet
// counts (entry P = 0, so that the exclusive scan's entry P is the total),
// then each packet's open lanes at its scanned offset.
__global__ __launch_bounds__(kBlock) void k_open_count(const uint64_t* __restrict__ open_mask, int64_t P,
                                                      uint32_t* __restrict__ cnt)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= P) cnt[b] = b < P ? (uint32_t)__popcll(open_mask[b]) : 0u;
}

__global__ __launch_bounds__(64) void k_open_fill(const uint64_t* __restrict__ open_mask,
                                                 const uint32_t* __restrict__ off, int32_t* __restrict__ list)
{
    const int64_t b = blockIdx.x;
    const uint64_t m = open_mask[b];
    const int lane = __lane_id();
    if ((m >> lane) & 1ull) list[off[b] + (uint32_t)__popcll(m & lanemask_lt())] = (int32_t)(b * 64 + lane);
}

// The same compaction in three launches: per-tile exclusive scan of the
// packets' open counts (entry P = 0), the tile sums, then each packet's
// lanes at its offset (src: entries of a query list, else positions); the
// total lands in off[P].
__global__ __launch_bounds__(kBlock) void k_open_scan_local(const uint64_t* __restrict__ open_mask, int64_t P,
                                                           uint32_t* __restrict__ off, uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[16];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t t = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < P ? (uint32_t)__popcll(open_mask[base + j]) : 0u;
        t += v[j];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(t, sh, total);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j <= P) off[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// the tile totals of tiles [0, t) summed by the whole wave (a few hundred
// tiles at most for 10^8 points: cheaper than a launch that scans them)
__device__ __forceinline__ uint32_t tiles_before(const uint32_t* __restrict__ sums, int64_t t)
{
    uint32_t a = 0;
    for (int64_t j = __lane_id(); j < t; j += 64) a += sums[j];
    for (int o = 32; o > 0; o >>= 1) a += (uint32_t)__shfl_xor((int)a, o);
    return a;
}

// (each wave takes a run of consecutive packets, the grid at most 2048
// blocks: the lists are built every round, and a launch of one block per
// packet costs more than its work when few are open; sums holds the tile
// totals of k_open_scan_local: each wave adds up the ones before its run's
// first tile, then the tiles it crosses)
__global__ __launch_bounds__(kBlock) void k_open_fill_scan(uint64_t* __restrict__ open_mask, int64_t P,
                                                          uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ sums,
                                                          const int32_t* __restrict__ src, int32_t* __restrict__ dst)
{
    const int lane = __lane_id();
    const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
    const int64_t w0 = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    if (w0 == 0) {
        const uint32_t pre = tiles_before(sums, P / kScanTile);
        if (lane == 0) off[P] += pre;
    }
    const int64_t per = (P + waves - 1) / waves;
    const int64_t b0 = w0 * per, b1 = min(P, b0 + per);
    if (b0 >= b1) return;
    int64_t t = b0 / kScanTile;
    uint32_t pre = tiles_before(sums, t);
    for (int64_t b = b0; b < b1; ++b) {
        if (b / kScanTile != t) {
            pre += sums[t];
            ++t;
        }
        const uint64_t m = wave_first_u64(open_mask[b]);
        if (m == 0ull) continue;
        // read once: cleared, so that the search that writes the next mask
        // need only write the words of its packets
        if (lane == 0) open_mask[b] = 0ull;
        if (!((m >> lane) & 1ull)) continue;
        const uint32_t o = off[b] + pre + (uint32_t)__popcll(m & lanemask_lt());
        dst[o] = src != nullptr ? src[b * 64 + lane] : (int32_t)(b * 64 + lane);
    }
}

void open_compact(uint64_t* open_mask, int64_t P, uint32_t* off, uint32_t* sums, const int32_t* src,
                  int32_t* dst, hipStream_t s)
{
    const int64_t nb = (P + 1 + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_open_scan_local, dim3((unsigned)nb), dim3(kBlock), 0, s, open_mask, P, off, sums);
    const unsigned gf = (unsigned)std::min<int64_t>((P + kBlock / 64 - 1) / (kBlock / 64), 2048);
    hipLaunchKernelGGL(k_open_fill_scan, dim3(gf), dim3(kBlock), 0, s, open_mask, P, off, (const uint32_t*)sums,
                       src, dst);
}

// Start of a group refresh: m2 = ~0 for every group, dup_split = 0.
__global__ __launch_bounds__(kBlock) void k_dup_reset(Counters* __restrict__ C, DupState dup)
{
    if (C->done || C->dup_settled) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0) C->dup_split = 0;
    for (int64_t g = t0; g < dup.ngroups; g += stride) dup.m2[g] = 0xffffffffu;
}

// Per group, the smallest member id whose component differs from the
// representative's (labels are distinct before round 1); dup_split is set
// when any group spans >= 2 components.  m2 holds ~0 on entry.
__global__ __launch_bounds__(kBlock) void k_dup_m2(const int32_t* __restrict__ label,
                                                  const float4* __restrict__ pts,
                                                  Counters* __restrict__ C, DupState dup)
{
    if (C->done || C->dup_settled) return;
    bool split = false;
    // every lane walks the same number of trips, so the wave-wide steps below
    // see all 64 lanes
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x; t0 < dup.npoints; t0 += stride) {
        const int64_t t = t0 + threadIdx.x;
        bool valid = false;
        int32_t g = 0;
        uint32_t id = 0xffffffffu;
        if (t < dup.npoints) {
            const int32_t p = dup.list[t];
            const int32_t r = dup.rep[p];
            if (label[p] != label[r]) {
                valid = true;
                g = dup.gslot[r];
                id = __float_as_uint(pts[p].w);
            }
        }
        split |= valid;
        // lanes of one group (list entries of a group are mostly adjacent)
        // merge into one atomic; a lower value already there skips it
        const uint64_t vm = __ballot(valid);
        if (vm == 0ull) continue;
        const int32_t lg = __shfl(g, (int)__builtin_ctzll(vm));
        const bool grp = valid && g == lg;
        uint32_t v = grp ? id : 0xffffffffu;
        for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
        if (__lane_id() == (int)__builtin_ctzll(vm) && __atomic_load_n(&dup.m2[lg], __ATOMIC_RELAXED) > v)
            atomicMin(&dup.m2[lg], v);
        if (valid && !grp && __atomic_load_n(&dup.m2[g], __ATOMIC_RELAXED) > id) atomicMin(&dup.m2[g], id);
    }
    if (__ballot(split) != 0ull && __lane_id() == 0 && __atomic_load_n(&C->dup_split, __ATOMIC_RELAXED) == 0u)
        atomicOr(&C->dup_split, 1u);
}

// Cones of the group points.  A point whose location is shared with another
// component (before round 1: with any other point) has d2 = 0 to it; the zero
// vector lies in exactly the '+'-face cones, where (0, smallest such id)
// beats every other key.  The smallest group id of another component is the
// representative's id when p's component differs from the representative's,
// else m2.  Otherwise the '+' cones are the representative's own (kept at its
// search; no point at the shared location is a candidate).  The '-' cones are
// always the representative's: the zero vector is in no '-' cone.  first_corr
// / first_key (round 1) take the smallest key over the point's cones.
template <int D>
__global__ __launch_bounds__(kBlock) void k_dup_fill(int64_t n, const int32_t* __restrict__ label,
                                                    const float4* __restrict__ pts,
                                                    const int32_t* __restrict__ inv,
                                                    uint64_t* __restrict__ cand_key,
                                                    int32_t* __restrict__ cand_pos,
                                                    int32_t* __restrict__ revive,
                                                    int32_t* __restrict__ first_corr,
                                                    uint64_t* __restrict__ first_key,
                                                    const Counters* __restrict__ C, DupState dup)
{
    using CC = ConeCfg<D>;
    if (C->done || C->dup_settled) return;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < dup.npoints;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = dup.list[t];
        const int32_t r = dup.rep[p];
        const int64_t g = dup.gslot[r];
        const uint32_t m2 = dup.m2[g];
        const bool zero = m2 != 0xffffffffu;
        uint32_t z = 0;
        int32_t zpos = -1;
        if (zero) {
            z = label[p] != label[r] ? __float_as_uint(pts[r].w) : m2;
            zpos = inv[z];
        }
        uint64_t kmin = kEmptyKey;
        int32_t pmin = -1;
        // the representative's '-' cones are read, never written, here
        for (int c = 0; c < CC::K; ++c) {
            const int64_t o = (int64_t)c * n;
            uint64_t k;
            int32_t q;
            if (((c / CC::NS) & 1) == 0) {
                const int64_t b = (int64_t)((c / CC::NS / 2) * CC::NS + c % CC::NS) * dup.ngroups + g;
                k = zero ? (uint64_t)z : dup.bkey[b];
                q = zero ? zpos : dup.bpos[b];
                cand_key[o + p] = k;
                cand_pos[o + p] = q;
            } else {
                k = cand_key[o + r];
                q = cand_pos[o + r];
                if (p != r) {
                    cand_key[o + p] = k;
                    cand_pos[o + p] = q;
                }
            }
            if (k < kmin) { kmin = k; pmin = q; }
        }
        if (first_corr) {
            first_corr[p] = pmin >= 0 ? pmin : -2;
            first_key[p] = kmin;
        }
        if (revive && revive[p] == -2) revive[p] = -1;
    }
}

// ------------------------------------------------------- Boruvka rounds
__global__ __launch_bounds__(kBlock) void k_round_init(int64_t n, int32_t* __restrict__ label,
                                                      int32_t* __restrict__ label2,
                                                      int32_t* __restrict__ eu,
                                                      int32_t* __restrict__ corr,
                                                      uint64_t* __restrict__ corr_key,
                                                      uint64_t* __restrict__ comp1,
                                                      uint32_t* __restrict__ comp2, Counters* C,
                                                      uint32_t target, uint32_t max_rounds)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        C->edges = 0; C->research = 0; C->zero_or_empty = 0; C->unflat = 0;
        C->packets = C->nodes = C->leaves = C->leaf_points = C->pairs = 0;
        C->round = 0; C->done = C->invalid ? 1u : 0u; C->edges_prev = 0; C->dup_settled = 0; C->nsamples = 0;
        C->comp_queries = 0;
        C->deferred = 0;
        C->knn_round1 = 0;
        C->target = target; C->max_rounds = max_rounds;
    }
    if (i < kWorkKinds * kWorkSlots) C->work[i / kWorkSlots][i % kWorkSlots] = 0;
    if (i >= n) return;
    label[i] = (int32_t)i;          // K_EMST_initDisjointSet (EMSTOperators.h:330-343)
    label2[i] = (int32_t)i;
    eu[i] = -1;                     // no edge absorbed this root yet
    corr[i] = -1;                   // correspondenceMap reset (SolutionEMST.cu:464-466)
    corr_key[i] = kEmptyKey;
    comp1[i] = kEmptyKey;
    comp2[i] = 0xffffffffu;
}

// round 1 with its targets from k_knn_first (Counters::knn_round1)
__device__ __forceinline__ bool knn_round1(const Counters* C) { return C->round == 1u && C->knn_round1 != 0u; }

// Termination is tested before a round (SolutionEMSTOperators.cu:194-227 and
// the 100-round cap): the tree is complete or the cap is reached.  While
// done != 0 every round kernel returns at once, so rounds can be issued
// ahead (a captured graph of several rounds) and stop by themselves.
__global__ void k_round_begin(Counters* C)
{
    if (C->done) return;
    if (C->edges >= C->target || C->round >= C->max_rounds) {
        C->done = 1;
        return;
    }
    C->round += 1;
    C->edges_prev = C->edges;
    C->nroots = 0;
    C->deferred = 0;
}

// A round that added no edge leaves a forest with no outgoing candidate: done.
// A flatten that hit its step cap pauses the rounds until the host has
// finished it.  Coincident groups stop being refreshed once none spans two
// components (components only merge).  stamps (one-launch runs): the
// device's constant-rate clock at the end of each round that ran, at
// stamps[round].
__global__ void k_round_end(Counters* C, int grouped, unsigned long long* stamps, uint32_t cap)
{
    if (C->done) return;
    if (stamps != nullptr && C->round <= cap) stamps[C->round] = wall_clock64();
    if (C->unflat) {
        C->done = 2;
        return;
    }
    if (grouped && !C->dup_settled && C->dup_split == 0u) C->dup_settled = 1;
    if (C->edges == C->edges_prev) C->done = 1;
}

__global__ void k_stamp(unsigned long long* at)
{
    *at = wall_clock64();
}

// Points that must search again next round: a zero-distance cone or no cone
// at all (NodeEMST.h:1051-1077).
__global__ __launch_bounds__(kBlock) void k_flag_research(int K, int64_t n,
                                                         const uint64_t* __restrict__ cand_key,
                                                         int32_t* __restrict__ list, Counters* C)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    bool zero = false, any = false;
    for (int k = 0; k < K; ++k) {
        const uint64_t key = cand_key[(int64_t)k * n + p];
        if (key == kEmptyKey) continue;
        any = true;
        zero |= (key >> 32) == 0u;
    }
    if (zero || !any) {
        atomicAdd(&C->zero_or_empty, 1u);
        list[atomicAdd(&C->research, 1u)] = (int32_t)p;
    }
}

// Component minima: one atomic per run of equal labels in a wave
// (atomic_min_grouped).  In round 2 of the clustered 10^7 cloud, with
// components of about 3 points, per-lane atomics took 0.15 of the keep pass's
// 0.35 ms.
// The minima only decrease, so a value read (even a stale one) that is
// already <= the key makes the atomic redundant: late rounds, whose few large
// components would otherwise take one atomic per wave on the same address,
// mostly skip it.
__device__ __forceinline__ void atomic_min_if_lower(uint64_t* a, uint64_t key)
{
    if (__atomic_load_n(a, __ATOMIC_RELAXED) > key) atomicMin((unsigned long long*)a, (unsigned long long)key);
}

__device__ __forceinline__ void atomic_min_grouped(uint64_t* arr, int32_t lab, uint64_t key, bool valid)
{
    if (__ballot(valid) == 0ull) return;
    // runs of equal labels across consecutive lanes (the lanes hold
    // consecutive positions or queries, and a component's points lie close
    // together in position order): a segmented minimum from the right leaves
    // each run's minimum in its first lane, which alone updates the entry;
    // a lane past the run reads its own label, so runs never merge
    const int lane = __lane_id();
    const int32_t prev = __shfl_up(lab, 1);
    const uint64_t heads = __ballot(lane == 0 || prev != lab) | 1ull;
    const int seg = 63 - __builtin_clzll(heads & (~0ull >> (63 - lane)));
    uint64_t v = valid ? key : kEmptyKey;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int src = lane + d < 64 ? lane + d : lane;
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
        const int oseg = __shfl(seg, src);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        if (oseg == seg && o < v) v = o;
    }
    if (seg == lane && v != kEmptyKey) atomic_min_if_lower(&arr[lab], v);
}

// The same with one group, the first valid lane's label: lanes holding it
// merge into one atomic, the rest issue their own.  Fewer live registers than
// the segmented form: in the component search (80 VGPRs at 6 waves per SIMD)
// the segmented form spilled 48 instead of 20 bytes a lane, doubled the
// search's writes (0.50 -> 0.99 GB per EMST) and cost 0.12 ms.
__device__ __forceinline__ void atomic_min_leader(uint64_t* arr, int32_t lab, uint64_t key, bool valid)
{
    const uint64_t vm = __ballot(valid);
    if (vm == 0ull) return;
    const int leader = __builtin_ctzll(vm);
    const int32_t llab = __shfl(lab, leader);
    const bool grp = valid && lab == llab;
    const uint64_t gm = __ballot(grp);
    uint64_t v = grp ? key : kEmptyKey;
    if (__popcll(gm) > 1) {
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = (uint64_t)__shfl_xor((long long)v, off);
            v = o < v ? o : v;
        }
        if (__lane_id() == leader) atomic_min_if_lower(&arr[llab], v);
        if (valid && !grp) atomic_min_if_lower(&arr[lab], key);
    } else if (valid) {
        atomic_min_if_lower(&arr[lab], key);
    }
}

// Each point's cone candidates in key order, for the rounds (bitonic network
// over K padded to a power of two, all indices static so the keys stay in
// registers): rkey[p*K + r], rpos[p*K + r] = key and sorted position of the
// candidate of rank r; cursor[p] = 0.  Equal keys carry equal positions (a
// key holds the candidate's id; empty cones and the padding are (~0, -1)),
// so the position rides along as payload and ranks 0..K-1 are exactly the K
// cones.  Point-major rows keep a point's next ranks in one cache line for
// k_round_min; each row is written with 16-byte stores.  Points: [0, n) or
// qlist[0..nq); with C the kernel does nothing once the rounds are done or
// the groups have settled (it follows a group refresh in a replayed round).
template <int D>
__global__ __launch_bounds__(kBlock) void k_rank_cones(int64_t n, const uint64_t* __restrict__ cand_key,
                                                      const int32_t* __restrict__ cand_pos,
                                                      const int32_t* __restrict__ qlist, int64_t nq,
                                                      uint64_t* __restrict__ rkey, int32_t* __restrict__ rpos,
                                                      uint8_t* __restrict__ cursor,
                                                      const Counters* __restrict__ C)
{
    using CC = ConeCfg<D>;
    constexpr int P = CC::K <= 8 ? 8 : 32;
    static_assert(CC::K % 4 == 0, "rows are written in 16-byte pieces");
    if (C && (C->done || C->dup_settled)) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int64_t p = qlist ? (int64_t)qlist[t] : t;
    uint64_t key[P];
    int32_t pos[P];
#pragma unroll
    for (int c = 0; c < P; ++c) {
        key[c] = c < CC::K ? cand_key[(int64_t)c * n + p] : kEmptyKey;
        pos[c] = c < CC::K ? cand_pos[(int64_t)c * n + p] : -1;
    }
#pragma unroll
    for (int k = 2; k <= P; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    if ((key[i] > key[l]) == up && key[i] != key[l]) {
                        const uint64_t tk = key[i]; key[i] = key[l]; key[l] = tk;
                        const int32_t tp = pos[i]; pos[i] = pos[l]; pos[l] = tp;
                    }
                }
            }
    uint4* __restrict__ ko = reinterpret_cast<uint4*>(rkey + p * CC::K);
    uint4* __restrict__ po = reinterpret_cast<uint4*>(rpos + p * CC::K);
#pragma unroll
    for (int r = 0; r < CC::K; r += 2)
        ko[r / 2] = make_uint4((uint32_t)key[r], (uint32_t)(key[r] >> 32), (uint32_t)key[r + 1],
                               (uint32_t)(key[r + 1] >> 32));
#pragma unroll
    for (int r = 0; r < CC::K; r += 4)
        po[r / 4] = make_uint4((uint32_t)pos[r], (uint32_t)pos[r + 1], (uint32_t)pos[r + 2], (uint32_t)pos[r + 3]);
    cursor[p] = 0;
}

// find min 1 (NodeEMST.h:1303-1384) + first phase of the component minimum:
// per point, keep the previous target while it is outside the component,
// else the (d2, id)-smallest cone candidate with another label; then
// atomicMin of (d2 bits, min id) on the component root.
//
// The smallest outgoing candidate is the first, in key order, whose label
// differs: the cones are ranked once (k_rank_cones) and cursor[p] skips the
// ones already found inside the component, which never leave it (components
// only merge), so each candidate is looked at about once over all rounds.
__global__ __launch_bounds__(kBlock) void k_round_min(int K, int64_t n, const uint32_t* __restrict__ ids,
                                                     const uint64_t* __restrict__ rkey,
                                                     const int32_t* __restrict__ rpos,
                                                     uint8_t* __restrict__ cursor,
                                                     const int32_t* __restrict__ label,
                                                     int32_t* __restrict__ corr,
                                                     uint64_t* __restrict__ corr_key,
                                                     uint64_t* __restrict__ comp1, const Counters* C)
{
    if (C->done) return;
    {
        const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool inb = p < n;
        int32_t lp = 0;
        uint64_t key1 = kEmptyKey;
        bool valid = false;
        if (inb) {
            lp = label[p];
            int32_t c = corr[p];
            uint64_t ck = corr_key[p];
            // corr == -2: every cone candidate already inside the component;
            // since components only merge, the point never has an outgoing
            // candidate again (until a repeated cone search or a group
            // refresh gives it new candidates)
            if (c != -2 && !(c >= 0 && label[c] != lp)) {
                c = -1;
                ck = kEmptyKey;
                // kStep ranks per step, their loads independent of each other,
                // so a long run of candidates inside the component costs few
                // round trips to memory
                constexpr int kStep = 4;
                int r = cursor[p];
                bool more = r < K;
                const uint64_t* __restrict__ kr = rkey + p * (int64_t)K;
                const int32_t* __restrict__ pr = rpos + p * (int64_t)K;
                while (more) {
                    uint64_t key[kStep];
                    int32_t q[kStep], lq[kStep];
#pragma unroll
                    for (int i = 0; i < kStep; ++i) {
                        key[i] = kEmptyKey;
                        q[i] = -1;
                        if (r + i < K) {
                            key[i] = kr[r + i];
                            q[i] = pr[r + i];
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kStep; ++i) lq[i] = key[i] != kEmptyKey ? label[q[i]] : lp;
#pragma unroll
                    for (int i = 0; i < kStep; ++i) {
                        if (!more) continue;
                        if (key[i] == kEmptyKey) {      // ranked last: no outgoing candidate left
                            r = K;
                            more = false;
                        } else if (lq[i] != lp) {
                            ck = key[i];
                            c = q[i];
                            more = false;
                        } else {
                            ++r;
                        }
                    }
                    more = more && r < K;
                }
                cursor[p] = (uint8_t)r;
                corr[p] = c == -1 ? -2 : c;
                corr_key[p] = ck;
            }
            if (c >= 0) {
                const uint32_t pid = ids[p];
                const uint32_t qid = (uint32_t)ck;
                key1 = (ck & 0xffffffff00000000ull) | (uint64_t)min(pid, qid);
                valid = true;
            }
        }
        atomic_min_grouped(comp1, lp, key1, valid);
    }
}

// a member's edge key, high part: (d2, min id); the max id completes it
__device__ __forceinline__ uint64_t edge_key1(uint64_t ck, uint32_t pid)
{
    return (ck & 0xffffffff00000000ull) | (uint64_t)min(pid, (uint32_t)ck);
}

// second phase: among members matching (d2, min id), the smallest max id.
__global__ __launch_bounds__(kBlock) void k_round_comp2(int64_t n, const uint32_t* __restrict__ ids,
                                                       const int32_t* __restrict__ label,
                                                       const int32_t* __restrict__ corr,
                                                       const uint64_t* __restrict__ corr_key,
                                                       const uint64_t* __restrict__ comp1,
                                                       uint32_t* __restrict__ comp2, const Counters* C)
{
    if (C->done || knn_round1(C)) return;       // one member per component (see k_round_hook)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || corr[p] < 0) return;
    const int32_t lp = label[p];
    const uint64_t ck = corr_key[p];
    const uint32_t pid = ids[p];
    if (edge_key1(ck, pid) == comp1[lp]) atomicMin(&comp2[lp], max(pid, (uint32_t)ck));
}

// hook/union (EMSTOperators.h:1267-1303): roots read from the flattened
// labels, written into label2.  label2 is a parent forest whose chase from
// any point ends at its component root (k_round_init starts it as the
// identity; only roots are written here and the flatten only halves paths),
// so it needs no copy of the labels between rounds.
// The winner of a component (fixedMap, NodeEMST.h:947-949) is the member
// whose edge key (d2, min id, max id) equals the component's comp1 / comp2;
// the pair of ids makes it unique, so the test is made here, for p and for
// its target q, with no per-point flag between passes.
// In round 1 every component is one point, whose edge is its own: with the
// targets from k_knn_first, every point with a target wins, and a pair is
// mutual when the target's target is the point; the component minima are
// then not needed (k_comp_keep and k_round_comp2 skip that round).
// Non-mutual: root(p) -> root(q).  Mutual pair (q won its component with an
// edge back to p): the endpoint with the smaller id keeps its root.  Each
// hook absorbs exactly one root and a root is absorbed at most once over the
// run, so the edge is stored at the absorbed root's index (eu = -1 marks an
// empty slot); the edge count is reduced per block.
constexpr int kHookBlock = 1024;

__global__ __launch_bounds__(kHookBlock) void k_round_hook(int64_t n, const uint32_t* __restrict__ ids,
                                                          const int32_t* __restrict__ label,
                                                          int32_t* __restrict__ label2,
                                                          const int32_t* __restrict__ corr,
                                                          const uint64_t* __restrict__ corr_key,
                                                          const uint64_t* __restrict__ comp1,
                                                          const uint32_t* __restrict__ comp2,
                                                          Counters* C, int32_t* __restrict__ eu,
                                                          int32_t* __restrict__ ev,
                                                          float* __restrict__ ed2,
                                                          int32_t* __restrict__ eround,
                                                          const uint8_t* __restrict__ cut)
{
    if (C->done) return;
    const int round = (int)C->round;
    const bool first = knn_round1(C);
    __shared__ uint32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    uint32_t added = 0;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t q = p < n ? corr[p] : -1;
    if (q >= 0) {
        const int32_t rp = first ? (int32_t)p : label[p];
        const uint64_t ck = corr_key[p];
        const uint32_t pid = ids[p], qid = (uint32_t)ck;
        const uint64_t key1 = edge_key1(ck, pid);
        if (first || (key1 == comp1[rp] && max(pid, qid) == comp2[rp] && !(cut && round > 1 && cut[rp]))) {
            const int32_t rq = first ? q : label[q];
            bool mutual = false;            // q won its component with the edge back to p
            if (corr[q] == (int32_t)p) {
                mutual = true;
                if (!first) {
                    const uint64_t cq = corr_key[q];
                    const uint32_t qq = ids[q];
                    mutual = edge_key1(cq, qq) == comp1[rq] && max(qq, (uint32_t)cq) == comp2[rq];
                }
            }
            int32_t absorbed = -1;
            if (!mutual) {
                label2[rp] = rq;
                absorbed = rp;
            } else if (pid < qid) {
                label2[rq] = rp;
                absorbed = rq;
            }
            if (absorbed >= 0) {
                ++added;
                eu[absorbed] = (int32_t)min(pid, qid);
                ev[absorbed] = (int32_t)max(pid, qid);
                ed2[absorbed] = __uint_as_float((uint32_t)(key1 >> 32));
                eround[absorbed] = round;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) added += (uint32_t)__shfl_xor((int)added, off);
    if (__lane_id() == 0 && added) atomicAdd(&cnt, added);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(&C->edges, cnt);
}

// ---- per-round component search (NodeSpiralSearchMD::search,
// emst_coalition/include/NodeEMST.h:67-509; the reference's alternative to
// the round-1 cone search, SolutionEMST.h:55-56,187-194).  Every round, each
// point whose previous target has joined its component looks for its nearest
// point in another component, under (d2, id); a point whose target is still
// outside keeps it (:160-169).  The search is exact per point, so a
// component's smallest member result is its shortest outgoing edge, and the
// rounds build the same forests as the cone search.  Subtrees that lie in
// the point's own component are skipped (node component labels, rebuilt
// every round), which plays the part of the reference's permanent slab
// closures against same-component points (:293-318).

constexpr int32_t kNlabEmpty = -3;      // a node without points
constexpr int32_t kNlabMixed = -1;      // a node holding several components

// Node values of the hierarchy, bottom-up in one launch: one thread per leaf
// climbs from its leaf; at each parent the first child to arrive leaves its
// value and stops, the second combines both and goes on.  The arrival count
// and the first value share one 64-bit word per node (`visit`, updated only
// by compare-and-swap, so no fence is needed between them); every launch
// brings exactly two arrivals to every inner node, so the count's parity
// tells first from second without a reset between launches.
template <class Op, class Store>
__device__ __forceinline__ void climb(typename Op::T v, Store store, unsigned long long* __restrict__ vbase,
                                      int64_t i)
{
    while (i > 1) {
        const int64_t par = i >> 1;
        unsigned long long old = vbase[par];
        bool first;
        for (;;) {
            const uint32_t cnt = (uint32_t)(old >> 32);
            first = (cnt & 1u) == 0u;
            const typename Op::T nv = first ? v : Op::combine((typename Op::T)(uint32_t)old, v);
            const unsigned long long nw = ((unsigned long long)(cnt + 1u) << 32) | (uint32_t)nv;
            const unsigned long long got = atomicCAS(&vbase[par], old, nw);
            if (got == old) {
                v = nv;
                break;
            }
            old = got;
        }
        if (first) return;
        i = par;
        store(i, v);
    }
}

struct LabelOp {
    typedef int32_t T;
    __device__ static int32_t combine(int32_t a, int32_t b)
    {
        return a == kNlabEmpty ? b : (b == kNlabEmpty ? a : (a == b ? a : kNlabMixed));
    }
};
struct MinOp {
    typedef uint32_t T;
    __device__ static uint32_t combine(uint32_t a, uint32_t b) { return min(a, b); }
};

// component label of every node: the common label of its points.  A block
// whose kBlock leaf slots are one aligned subtree of a cloud (every block of a
// cloud with at least kBlock slots) combines the subtree's 8 lower levels in
// LDS, and only the subtree's root climbs with atomics: the climb of every
// leaf through those levels was a chain of dependent atomics per level.  The
// arrival counts of the nodes inside such a subtree are then left alone,
// which keeps their parity for the other climbs (k_nmin_tree).
__global__ __launch_bounds__(kBlock) void k_nlab_tree(const int32_t* __restrict__ label,
                                                     const Cloud* __restrict__ cl, int ncl, int64_t nleaves,
                                                     LeafTable lt, Node* __restrict__ nodes,
                                                     unsigned long long* __restrict__ visit, const Counters* C)
{
    if (C->done || knn_round1(C)) return;       // round 1 searches nothing
    __shared__ int32_t sub_lab[kBlock];
    const int64_t g0 = (int64_t)blockIdx.x * kBlock;
    const int cb = find_cloud_leaf(cl, ncl, g0);
    const int64_t jb = g0 - cl[cb].lbase;
    const bool subtree = (jb & (kBlock - 1)) == 0 && jb + kBlock <= cl[cb].L;
    const int64_t g = g0 + threadIdx.x;
    if (!subtree && g >= nleaves) return;
    const int c = subtree ? cb : find_cloud_leaf(cl, ncl, g);
    const int64_t j = g - cl[c].lbase;
    if (!subtree && j >= cl[c].L) return;   // past the cloud's slots (one-launch capacity)
    int64_t s0, s1;
    leaf_range(lt, cl[c].lfirst, cl[c].nl, cl[c].pos + cl[c].n, j, s0, s1);
    // the leaf's common label (kNlabMixed when two differ), its loads
    // independent of each other (a serial combine waited on each load: the
    // kernel took 0.12 ms per round at 10^7 points)
    int32_t lab = kNlabEmpty;
    if (s1 > s0) {
        lab = label[s0];
        bool diff = false;
        int64_t s = s0 + 1;
        for (; s + 8 <= s1; s += 8) {
            int32_t v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = label[s + k];
#pragma unroll
            for (int k = 0; k < 8; ++k) diff |= v[k] != lab;
        }
        for (; s < s1; ++s) diff |= label[s] != lab;
        if (diff) lab = kNlabMixed;
    }
    Node* base = nodes + 2 * cl[c].lbase;
    auto store = [&](int64_t i, int32_t v) { base[i].pad[1] = __int_as_float(v); };
    const int64_t L = cl[c].L;
    store(L + j, lab);
    if (!subtree) {
        climb<LabelOp>(lab, store, visit + 2 * cl[c].lbase, L + j);
        return;
    }
    // the subtree's levels: at level k (leaves at 0), thread t < kBlock >> k
    // holds node ((L + jb) >> k) + t
    const int t = (int)threadIdx.x;
    sub_lab[t] = lab;
    int32_t v = lab;
    for (int k = 1, w = kBlock / 2; w >= 1; ++k, w >>= 1) {
        __syncthreads();
        if (t < w) v = LabelOp::combine(sub_lab[2 * t], sub_lab[2 * t + 1]);
        __syncthreads();
        if (t < w) {
            sub_lab[t] = v;
            store(((L + jb) >> k) + t, v);
        }
    }
    if (t == 0) climb<LabelOp>(v, store, visit + 2 * cl[c].lbase, (L + jb) / kBlock);
}

// smallest point id under every node (static per build): with ties on d2
// broken by id, a node at the bound can only matter when it holds a smaller
// id than the best so far (runs of coincident points)
__global__ __launch_bounds__(kBlock) void k_nmin_tree(const float4* __restrict__ pts,
                                                     const Cloud* __restrict__ cl, int ncl, int64_t nleaves,
                                                     LeafTable lt, Node* __restrict__ nodes,
                                                     unsigned long long* __restrict__ visit)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nleaves) return;
    const int c = find_cloud_leaf(cl, ncl, g);
    const int64_t j = g - cl[c].lbase;
    if (j >= cl[c].L) return;               // past the cloud's slots (one-launch capacity)
    int64_t s0, s1;
    leaf_range(lt, cl[c].lfirst, cl[c].nl, cl[c].pos + cl[c].n, j, s0, s1);
    uint32_t m = 0xffffffffu;
    int64_t s = s0;
    for (; s + 4 <= s1; s += 4) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = __float_as_uint(pts[s + j].w);
        m = min(m, min(min(v[0], v[1]), min(v[2], v[3])));
    }
    for (; s < s1; ++s) m = min(m, __float_as_uint(pts[s].w));
    Node* base = nodes + 2 * cl[c].lbase;
    auto store = [&](int64_t i, uint32_t v) { base[i].pad[0] = __uint_as_float(v); };
    store(cl[c].L + j, m);
    climb<MinOp>(m, store, visit + 2 * cl[c].lbase, cl[c].L + j);
}

// Per point, corr / corr_key between rounds:
//   corr >= 0   the exact nearest point outside the component at its last
//               search, key (d2, id); kept while it stays outside
//               (NodeEMST.h:160-169);
//   corr == -1  no exact result; the high word of corr_key is a lower bound
//               on the d2 of the nearest outside point (all ones: none);
//   corr == -2  nothing outside the component in its cloud (final).
// Keep-previous: the component minimum of the kept targets.
template <int K>
__global__ __launch_bounds__(kBlock) void k_comp_keep(int64_t n, const uint32_t* __restrict__ ids,
                                                     const int32_t* __restrict__ label,
                                                     int32_t* __restrict__ corr,
                                                     uint64_t* __restrict__ corr_key,
                                                     uint64_t* __restrict__ comp1, NbrList nb, const Counters* C)
{
    if (C->done || knn_round1(C)) return;       // every target is outside and kept
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    int32_t lp = 0;
    uint64_t key1 = kEmptyKey;
    if (p < n) {
        lp = label[p];
        int32_t c = corr[p];
        const int j0 = nb.pos != nullptr ? (int)nb.cur[p] : K;
        bool outside = c >= 0 && label[c] != lp;
        uint64_t ck = 0;
        bool have_ck = false;
        // a target that joined the component: the next listed point outside
        // it (components only grow, so the listed points before it are inside
        // for good); the list used up, its last key bounds the search.  The
        // remaining entries, then their labels, load at once: a scan entry by
        // entry waited on two dependent loads per entry (the kernel took up to
        // 0.36 ms per round at 10^7 points, when every target has just joined).
        if (c >= 0 && !outside && j0 < K) {
            int32_t e[K], le[K];
#pragma unroll
            for (int k = 1; k < K; ++k) e[k] = k > j0 ? nb.pos[(int64_t)k * n + p] : -1;
            // (unconditional loads, an absent entry reading p's own label, so
            // that they issue together rather than one branch and wait each)
#pragma unroll
            for (int k = 1; k < K; ++k) le[k] = label[e[k] >= 0 ? e[k] : (int32_t)p];
            int j = K;
            int32_t f = -1;
#pragma unroll
            for (int k = K - 1; k >= 1; --k)
                if (k > j0 && (e[k] < 0 || le[k] != lp)) {
                    j = k;
                    f = e[k];
                }
            if (j == K) {
                c = -1;
                ck = (nb.key[(int64_t)(K - 1) * n + p] & 0xffffffff00000000ull) | 0xffffffffull;
            } else if (f < 0) {
                c = -2;             // the cloud's other points are all listed, all inside
                ck = kEmptyKey;
            } else {
                c = f;
                ck = nb.key[(int64_t)j * n + p];
                have_ck = true;
                outside = true;
            }
            nb.cur[p] = (uint8_t)j;
            corr[p] = c;
            corr_key[p] = ck;
        }
        if (outside) {
            if (!have_ck) ck = corr_key[p];
            const uint32_t pid = ids[p], qid = (uint32_t)ck;
            key1 = (ck & 0xffffffff00000000ull) | (uint64_t)min(pid, qid);
            valid = true;
        }
    }
    atomic_min_grouped(comp1, lp, key1, valid);
}

// The points that must search, flagged per 64-position packet: those without
// a kept target whose lower bound does not already exceed the component's
// kept minimum (such a point cannot hold the component's shortest edge this
// round; it keeps its bound for the next), and that do not sit at the same
// location as the position before them in the same component.  Such a point
// is dominated: equal keys sort stably, so the point before it has a smaller
// id, the same nearest outside point at the same d2, and so a key (d2, min
// id, max id) at least as small; the first of the run searches for both (the
// coincident corners of a clamped cloud are runs of 10^4-10^5 points).
__global__ __launch_bounds__(kBlock) void k_comp_prep(int64_t n, const float4* __restrict__ pts,
                                                     const int32_t* __restrict__ label,
                                                     int32_t* __restrict__ corr,
                                                     const uint64_t* __restrict__ corr_key,
                                                     const uint64_t* __restrict__ comp1,
                                                     uint64_t* __restrict__ open_mask, const Counters* C)
{
    if (C->done) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    if (p < n && !knn_round1(C)) {          // else every point holds its nearest neighbour
        const int32_t lp = label[p];
        const int32_t c = corr[p];
        if (c == -1 || (c >= 0 && label[c] == lp)) {
            const uint32_t lb = (uint32_t)(corr_key[p] >> 32);
            const int32_t lbb = lb == 0xffffffffu ? -1 : (int32_t)lb;
            need = lbb <= key_bound(comp1[lp]);
            if (need && p > 0 && label[p - 1] == lp && same_xyz(pts[p - 1], pts[p])) need = false;
            if (!need) corr[p] = -1;    // the absorbed target's d2 stays as the bound
        }
    }
    const uint64_t m = __ballot(need);
    if (__lane_id() == 0 && p < n) open_mask[p / 64] = m;
}

// ---- learned seed of the component search.  A component's size and box
// live at its root: single points at the start, and after every round each
// root absorbed in it adds its own to its new root (one atomic per merged
// component and box side).
__global__ __launch_bounds__(kBlock) void k_comp_seed_init(int64_t n, const float4* __restrict__ pts,
                                                          int32_t* __restrict__ csize, uint32_t* __restrict__ cbox)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 v = pts[p];
    csize[p] = 1;
    uint32_t* b = cbox + 6 * p;
    b[0] = b[3] = f2ord(v.x);
    b[1] = b[4] = f2ord(v.y);
    b[2] = b[5] = f2ord(v.z);
}

__global__ __launch_bounds__(kBlock) void k_comp_merge(int64_t n, const int32_t* __restrict__ label,
                                                      const int32_t* __restrict__ eu,
                                                      const int32_t* __restrict__ eround,
                                                      int32_t* __restrict__ csize, uint32_t* __restrict__ cbox,
                                                      const Counters* C)
{
    if (C->done) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int prev = (int)C->round - 1;
    if (prev < 1 || eu[p] < 0 || eround[p] != prev) return;
    const int64_t r = label[p];
    atomicAdd(&csize[r], csize[p]);
    for (int k = 0; k < 3; ++k) {
        atomicMin(&cbox[6 * r + k], cbox[6 * p + k]);
        atomicMax(&cbox[6 * r + 3 + k], cbox[6 * p + 3 + k]);
    }
}

// Features of every searching point (list order), free of the cloud's scale:
// lengths are taken relative to s0, the mean spacing of the point's leaf cell
// (LS = log s0).  The round, the component's size, the point's lower bound,
// the component's box extent and mean spacing, the component's kept minimum
// (an upper bound on its edge), and the density one and three levels up.
// The model seeds rounds from kLearnedFrom on: the first rounds' searches are
// nearest-neighbour searches of single points and tiny components, for which
// the plain seed is as good and the model's cost would not pay.
// nq_keep[0]: the count of points with features (every round when logging,
// else from kLearnedFrom on), for the sample log; nq_keep[1]: the count the
// model runs on (0 before kLearnedFrom or without a model).
constexpr uint32_t kLearnedFrom = 3;

__global__ __launch_bounds__(kBlock) void k_comp_features(int log_all, int has_model, int dim, const float4* __restrict__ pts,
                                                         const Node* __restrict__ nodes,
                                                         const Cloud* __restrict__ clouds, int ncl, LeafTable lt,
                                                         const int32_t* __restrict__ qlist,
                                                         const uint32_t* __restrict__ nq_dev,
                                                         const int32_t* __restrict__ label,
                                                         const uint64_t* __restrict__ corr_key,
                                                         const uint64_t* __restrict__ comp1,
                                                         const int32_t* __restrict__ csize,
                                                         const uint32_t* __restrict__ cbox, const Counters* C,
                                                         float* __restrict__ X, float* __restrict__ LS,
                                                         uint32_t* __restrict__ nq_keep)
{
    if (C->done) return;
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool learned_round = has_model && C->round >= kLearnedFrom;
    const uint32_t nq = (log_all || learned_round) ? *nq_dev : 0u;
    if (qi == 0) {
        nq_keep[0] = log_all ? nq : 0u;
        nq_keep[1] = learned_round ? nq : 0u;
    }
    if (qi >= (int64_t)nq) return;
    const int64_t p = qlist[qi];
    const int c = find_cloud(clouds, ncl, p);
    const int64_t L = clouds[c].L, lfirst = clouds[c].lfirst, nl = clouds[c].nl;
    const int64_t cend = clouds[c].pos + clouds[c].n;
    const Node* __restrict__ tree = nodes + 2 * clouds[c].lbase;
    int64_t node = L + (lt.of[p] - lfirst);
    float sp[4];
    for (int k = 0; k < 4; ++k) {
        const Node nd = tree[node];
        int64_t first = node, last = node;
        while (first < L) { first = 2 * first; last = 2 * last + 1; }
        int64_t a0, a1, b0, b1;
        leaf_range(lt, lfirst, nl, cend, first - L, a0, a1);
        leaf_range(lt, lfirst, nl, cend, last - L, b0, b1);
        const float cnt = (float)max<int64_t>(b1 - a0, 1);
        float vol = 1.f, ext = 0.f;
        for (int d = 0; d < dim; ++d) {
            const float e = fmaxf(nd.hi[d] - nd.lo[d], 0.f);
            vol *= e;
            ext = fmaxf(ext, e);
        }
        const float sv = vol > 0.f ? powf(vol / cnt, 1.f / dim) : ext / cnt;
        sp[k] = fmaxf(sv, 1e-6f);
        if (node > 1) node >>= 1;
    }
    const float s0 = sp[0], lo = 1e-3f * s0;
    const int32_t lab = label[p];
    const uint32_t lbh = (uint32_t)(corr_key[p] >> 32);
    const float lb = lbh == 0xffffffffu ? 0.f : sqrtf(__uint_as_float(lbh));
    const uint64_t kept = comp1[lab];
    const float kd = kept == kEmptyKey ? 1e3f * s0 : sqrtf(__uint_as_float((uint32_t)(kept >> 32)));
    const uint32_t* b = cbox + 6 * (int64_t)lab;
    float ext = 0.f, vol = 1.f;
    for (int d = 0; d < dim; ++d) {
        const float e = fmaxf(ord2f(b[3 + d]) - ord2f(b[d]), 0.f);
        ext = fmaxf(ext, e);
        vol *= e;
    }
    const float size = (float)max(csize[lab], 1);
    const float spc = vol > 0.f ? powf(vol / size, 1.f / dim) : ext / size;
    float* x = X + qi * 8;
    x[0] = logf((float)max(C->round, 1u));
    x[1] = logf(size);
    x[2] = logf(fmaxf(lb, lo) / s0);
    x[3] = logf(fmaxf(ext, lo) / s0);
    x[4] = logf(fmaxf(spc, lo) / s0);
    x[5] = logf(fmaxf(kd, lo) / s0);
    x[6] = logf(sp[1] / sp[0]);
    x[7] = logf(sp[3] / sp[0]);
    LS[qi] = logf(s0);
}

// seed radius per sorted position: s0 * exp(model output + offset)
__global__ __launch_bounds__(kBlock) void k_comp_seed(const int32_t* __restrict__ qlist,
                                                     const uint32_t* __restrict__ nq_dev,
                                                     const float* __restrict__ LS, const float* __restrict__ Y,
                                                     float* __restrict__ R, const Counters* C)
{
    if (C->done) return;
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= (int64_t)*nq_dev) return;
    R[qlist[qi]] = expf(LS[qi] + fminf(Y[qi], 20.f));
}

// (features, target) of every stride-th searching point, stride chosen per
// round so that every round logs about `per_round` pairs (the late rounds,
// with few searching points, weigh as much as the first), logged from the
// plain method's rounds to train the model.  Target: log over s0 of the
// radius the point had to search, min(D, d): D = its component's edge length,
// d = its nearest outside point's distance when the search found it exactly
// (beyond D the component's edge stops the point).  Components without an
// outgoing edge are skipped.
__global__ __launch_bounds__(kBlock) void k_comp_samples(const int32_t* __restrict__ qlist,
                                                        const uint32_t* __restrict__ nq_dev,
                                                        const float* __restrict__ X, const float* __restrict__ LS,
                                                        const int32_t* __restrict__ label,
                                                        const int32_t* __restrict__ corr,
                                                        const uint64_t* __restrict__ corr_key,
                                                        const uint64_t* __restrict__ comp1, int64_t per_round,
                                                        int64_t cap, float* __restrict__ samples, Counters* C)
{
    if (C->done) return;
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nq = *nq_dev;
    const int64_t stride = max<int64_t>(1, (nq + per_round - 1) / per_round);
    if (qi >= nq || qi % stride != 0) return;
    const int64_t p = qlist[qi];
    const uint64_t k = comp1[label[p]];
    if (k == kEmptyKey) return;
    uint32_t db = (uint32_t)(k >> 32);
    if (corr[p] >= 0) db = min(db, (uint32_t)(corr_key[p] >> 32));
    const uint32_t slot = atomicAdd(&C->nsamples, 1u);
    if ((int64_t)slot >= cap) return;
    float* o = samples + 9 * (int64_t)slot;
    for (int f = 0; f < 8; ++f) o[f] = X[qi * 8 + f];
    const float s0 = expf(LS[qi]);
    const float d = sqrtf(__uint_as_float(db));
    o[8] = logf(fmaxf(d, 1e-3f * s0)) - LS[qi];
}

// LDS lists of the leaf gatherings (k_comp_search, k_knn_first): node ids of
// the work list, and gathered leaves as (box d2 bits << 32) | leaf
constexpr int kGatherWork = 512;
constexpr int kGatherLeaves = 128;
// Packets whose gathering outgrows the LDS lists hand their open queries to
// the per-query pass (k_comp_search_deferred) rather than walk the tree with
// all of them: in 3D up to kDeferCap queries a round (a walk over the union
// of 64 neighbourhoods is the tail of the large dispatches, and the per-query
// pass takes about as long for 64k queries as one such walk), in 2D only in
// dispatches of at most kDeferQueries queries (1024 packets, one per SIMD of
// the chip; in the large 2D dispatches the walks cost less than the pass)
constexpr int64_t kDeferQueries = 65536;
constexpr uint32_t kDeferCap = 65536;
// the resumed dispatch of the component search runs on at most this many
// one-packet blocks, looping over its packets: its queries are few (most
// close in the first pass), and a grid of one block per packet of the whole
// query list cost 34 us per round at 10^7 points while idle
constexpr int64_t kResumeBlocks = 4096;
// the first dispatch: at most this many blocks (far more than the chip holds,
// so the hardware still balances the packets); the compaction clears the
// mask, so a block need only write the words of the packets it searched
constexpr int64_t kSearchBlocks = 65536;
// Queries per packet.  A dispatch of few queries (rounds where the searching
// points are scattered) has fewer 64-query packets than the chip holds waves,
// and each packet walks the union of its queries' neighbourhoods: it takes
// kSmallPacket queries per packet instead, so that more, narrower packets run
// at once.
constexpr int64_t kSmallDispatch = 786432;
constexpr int kSmallPacket = 8;
// the neighbour lists' gatherings (k_knn_first): their radius is the packet's
// largest K-th distance after its own leaves, so they reach further
constexpr int kKnnGatherWork = 1024;
constexpr int kKnnGatherLeaves = 256;

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// A gathering starts from every node of depth min(6, leaf depth), 64 at most,
// one per lane: the levels above would each take a dependent step for a few
// nodes.  Returns the count written to work[].
__device__ __forceinline__ int gather_start(int64_t L, int32_t* work)
{
    int d0 = 0;
    while (d0 < 6 && ((int64_t)2 << d0) <= L) ++d0;
    const int cnt = 1 << d0;
    const int lane = __lane_id();
    if (lane < cnt) work[lane] = cnt + lane;
    __builtin_amdgcn_wave_barrier();
    return cnt;
}

// The gathering's next nodes from the inner nodes that passed: each hands on
// its 4 grandchildren (one 128-byte line of the heap), or its 2 children when
// those are leaves.  Every gathering test (box distance, far side, component
// label, empty) is monotone under box inclusion, so a grandchild whose parent
// would have failed fails too: the same nodes pass as level by level, in half
// the dependent steps.  Writes the entries at work[tail ..] and their count to
// add; false (nothing written) when they would pass cap.
__device__ __forceinline__ bool push_inner(bool pass, int nd, int64_t L, int32_t* work, int tail, int cap,
                                           int& add)
{
    const bool deep = pass && 2 * (int64_t)nd < L;     // its children are inner nodes
    const uint64_t m4 = __ballot(deep), m2 = __ballot(pass && !deep);
    add = 4 * __popcll(m4) + 2 * __popcll(m2);
    if (tail + add > cap) return false;
    const uint64_t lt = lanemask_lt();
    const int r = tail + 4 * __popcll(m4 & lt) + 2 * __popcll(m2 & lt);
    if (deep) {
#pragma unroll
        for (int k = 0; k < 4; ++k) work[r + k] = 4 * nd + k;
    } else if (pass) {
        work[r] = 2 * nd;
        work[r + 1] = 2 * nd + 1;
    }
    return true;
}

// ascending bitonic sort of one 64-bit key per lane across the wave
__device__ __forceinline__ uint64_t wave_sort_u64(uint64_t e)
{
    const int lane = __lane_id();
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t o = shfl_xor_u64(e, j);
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
            e = keep_min ? (o < e ? o : e) : (o > e ? o : e);
        }
    }
    return e;
}

// nearest point in another component, for the points of qlist (count in
// nq_dev): packet traversal as in the cone search, one best per query, by
// iterative deepening from the seed radius R_0 (capped passes R_k = 8^(k/2) R_0,
// then one uncapped).  Points below the point's lower bound (its previous
// target's d2) were in the component at that search and still are, so they
// are skipped.  The component's shortest edge found so far (comp1, read at
// each pass) also bounds the search: a point whose nearest outside point lies
// beyond it cannot hold the component's edge, and it stops with that d2 as
// its lower bound (corr = -1).  Exact results enter comp1 as they close.
template <int D, int BK>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_comp_search(
    const float4* __restrict__ pts, const Node* __restrict__ nodes,
    const Cloud* __restrict__ clouds, int ncl, int bucket_rt, LeafTable lt, const int32_t* __restrict__ qlist,
    const uint32_t* __restrict__ nq_dev, const int32_t* __restrict__ label, const int32_t* __restrict__ inv,
    float seed_scale, int32_t* __restrict__ corr, uint64_t* __restrict__ corr_key,
    uint64_t* __restrict__ comp1, int pass_lo, int pass_hi,
    uint64_t* __restrict__ open_out, const float* __restrict__ seed_r, uint64_t* __restrict__ defer, Counters* C,
    int collect)
{
    // the leaf size: a compile-time constant for the default sizes
    const int bucket = BK ? BK : bucket_rt;
    if (C->done) return;
    const int64_t nq = *nq_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&C->comp_queries, (unsigned long long)nq);
    __shared__ unsigned long long best[64];
    __shared__ int32_t stk_id[64];
    __shared__ int32_t stk_ver[64];
    __shared__ unsigned long long stk_mask[64];
    __shared__ int32_t work[kGatherWork];
    __shared__ unsigned long long gathered[kGatherLeaves];
    // one packet per block; a grid smaller than the packets (the resumed
    // dispatch, whose queries are few but unknown to the host) loops
    const int qpp = nq <= kSmallDispatch ? kSmallPacket : 64;
    for (int64_t pk = blockIdx.x;; pk += gridDim.x) {
        if (pk * qpp >= nq) {
            if (open_out != nullptr && threadIdx.x == 0 && pk == (int64_t)blockIdx.x) open_out[pk] = 0ull;
            break;
        }
        const int lane = __lane_id();
        const int64_t qi = pk * qpp + lane;
        const bool active = lane < qpp && qi < nq;
        const int64_t p = active ? (int64_t)qlist[qi] : 0;
        const float4 me = pts[p];
        const float px = me.x, py = me.y, pz = me.z;
        const int32_t mylab = active ? label[p] : -1;
        const uint64_t prev = active ? corr_key[p] : kEmptyKey;
        best[lane] = kEmptyKey;
        const int my_cl = active ? find_cloud(clouds, ncl, p) : -1;
        const uint32_t pid = __float_as_uint(me.w);
        bool bounded = false, left_open = false, deferred = false;
        int32_t lbnd = 0;
        uint32_t st_nodes = 0, st_leaves = 0, st_walks = 0;   // work counters of this packet
        uint64_t st_pairs = 0;
        uint64_t pending = __ballot(active);
        while (pending) {
            const int cl = __builtin_amdgcn_readfirstlane(__shfl(my_cl, (int)__builtin_ctzll(pending)));
            const bool mine = my_cl == cl;
            pending &= ~__ballot(mine);
            const int64_t cend = clouds[cl].pos + clouds[cl].n;
            const int64_t L = clouds[cl].L, lfirst = clouds[cl].lfirst, nl = clouds[cl].nl;
            const Node* __restrict__ tree = nodes + 2 * clouds[cl].lbase;
            // seed: twice the previous target's distance, at least seed_scale
            // times the mean spacing of the point's leaf cell
            float r = 1.f;
            int32_t inner = -1;
            if (mine) {
                const int64_t lf = lt.of[p];
                const Node own = tree[L + (lf - lfirst)];
                const float cnt = (float)(lt.start[lf + 1] - lt.start[lf]);
                float vol = 1.f, ext = 0.f;
                for (int k = 0; k < D; ++k) {
                    const float e = own.hi[k] - own.lo[k];
                    vol *= e;
                    ext = fmaxf(ext, e);
                }
                float sp = vol > 0.f ? powf(vol / cnt, 1.f / D) : ext / cnt;
                if (!positive_finite(sp)) sp = 1e-6f;
                const bool has_lb = (uint32_t)(prev >> 32) != 0xffffffffu;
                const uint32_t dbits = (uint32_t)(prev >> 32);
                if (has_lb) inner = (int32_t)dbits - 1;
                r = seed_scale * sp;
                if (has_lb) r = fmaxf(r, 2.f * sqrtf(__uint_as_float(dbits)));
                // learned: the predicted length of the component's edge, so
                // that the first pass can already find it
                if (seed_r != nullptr && C->round >= kLearnedFrom) {
                    r = seed_r[p];
                    if (has_lb) r = fmaxf(r, sqrtf(__uint_as_float(dbits)));
                }
                if (!positive_finite(r)) r = 1e-6f;
            }
            float r2 = r * r;
            // a launch that resumes at pass_lo: the state the earlier passes left
            for (int k = 0; k < pass_lo; ++k) {
                inner = max(inner, (int32_t)__float_as_uint(r2));
                r2 *= 8.f;
            }
            bool open = mine;
            float cx = mine ? px : 0.f, cy = mine ? py : 0.f, cz = mine ? pz : 0.f;
            float na = mine ? 1.f : 0.f;
            for (int off = 32; off > 0; off >>= 1) {
                cx += __shfl_xor(cx, off);
                cy += __shfl_xor(cy, off);
                cz += __shfl_xor(cz, off);
                na += __shfl_xor(na, off);
            }
            cx /= na; cy /= na; cz /= na;
            auto cdist = [&](const Node& x) {
                const float dxx = fmaxf(fmaxf(x.lo[0] - cx, cx - x.hi[0]), 0.f);
                const float dyy = fmaxf(fmaxf(x.lo[1] - cy, cy - x.hi[1]), 0.f);
                const float dzz = fmaxf(fmaxf(x.lo[2] - cz, cz - x.hi[2]), 0.f);
                return dxx * dxx + dyy * dyy + dzz * dzz;
            };
            // a box is needed when it holds another component, its nearest point
            // is within the limit and may beat the best so far under (d2, id)
            // (at equal d2 only with a smaller id), and its farthest point lies
            // beyond `inner`
            auto needed = [&](const Node& nd, int32_t nl, uint32_t nm, int32_t lim, uint64_t bk) {
                if (!open || nl == mylab || nl == kNlabEmpty) return false;
                float m[3], u[3];
                const float lo[3] = {nd.lo[0] - px, nd.lo[1] - py, nd.lo[2] - pz};
                const float hi[3] = {nd.hi[0] - px, nd.hi[1] - py, nd.hi[2] - pz};
    #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    m[k] = fmaxf(fmaxf(lo[k], -hi[k]), 0.f);
                    u[k] = fmaxf(-lo[k], hi[k]);
                }
                const float lb = D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]);
                const float far = D == 3 ? d2_of(u[0], u[1], u[2]) : d2_of(u[0], u[1]);
                const uint32_t lbb = __float_as_uint(lb);
                return (int32_t)lbb <= lim && (((uint64_t)lbb << 32) | nm) < bk &&
                       (int32_t)__float_as_uint(far) > inner;
            };
            for (int pass = pass_lo; pass <= pass_hi; ++pass) {
                const bool last = pass == kMaxCappedPasses;
                const int32_t cbb = open ? key_bound(comp1[mylab]) : 0x7fffffff;
                // a point whose component already has an edge (a kept target, or
                // a member's result) searches straight out to it in one pass
                // instead of deepening by 4x per pass
                const int32_t cap = last || cbb != 0x7fffffff ? 0x7fffffff : (int32_t)__float_as_uint(r2);
                // nothing outside the component lies within `inner`: with the
                // component's edge at most that long, the point cannot hold it
                if (open && cbb <= inner) {
                    open = false;
                    bounded = true;
                    lbnd = inner;
                }
                if (__ballot(open) == 0ull) break;
                const int32_t lim = min(cap, cbb);
                uint64_t bk = best[lane];
                int32_t bnd = min(key_bound(bk), lim);
                int sp = 0;
                int ver = 0;
                // first pass (3D): the leaf of the first open lane goes first, so
                // that the packet's lanes enter the tree with bounds from their
                // own neighbourhood; the traversal skips it when it comes again
                // (clustered 3D -3.4%; in 2D +1%, so not there)
                int pre = -1;
                bool pre_done = false;
                if (D == 3 && pass == 0) {
                    const int fl = (int)__builtin_ctzll(__ballot(open));
                    const int myleaf = (int)(L + (mine ? lt.of[p] - lfirst : 0));
                    pre = __builtin_amdgcn_readfirstlane(__shfl(myleaf, fl));
                }
                // a leaf: one (query, leaf point) pair per lane; the query's state by
                // LDS permute (a v_readlane broadcast when 64 / bucket = 1, as in
                // the cone search, measured 7% slower here, and four packets per
                // block 6% slower)
                auto visit_leaf = [&](int node, bool need, uint64_t needm) {
                    int64_t s0, s1;
                    leaf_range(lt, lfirst, nl, cend, node - L, s0, s1);
                    const int cnt = (int)(s1 - s0);
                    ++st_leaves;
                    st_pairs += (uint64_t)cnt * (uint64_t)__popcll(needm);
                    const int j = lane & (bucket - 1);
                    const int qslot = lane / bucket;
                    const int qpc = 64 / bucket;
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    int32_t vlab = -1;
                    if (j < cnt) {
                        v = pts[s0 + j];
                        vlab = label[s0 + j];
                    }
                    uint64_t rem = needm;
                    while (rem) {
                        int myq = -1;
                        for (int k = 0; k < qpc; ++k) {
                            const int l = rem ? (int)__builtin_ctzll(rem) : -1;
                            if (rem) rem &= rem - 1;
                            if (k == qslot) myq = l;
                        }
                        const int src = myq >= 0 ? myq : 0;
                        const float qx = __shfl(px, src), qy = __shfl(py, src), qz = __shfl(pz, src);
                        const int32_t qb = __shfl(bnd, src), qin = __shfl(inner, src), ql = __shfl(mylab, src);
                        const float dx = v.x - qx, dy = v.y - qy, dz = v.z - qz;
                        const float d2 = D == 3 ? d2_of(dx, dy, dz) : d2_of(dx, dy);
                        const int32_t db = (int32_t)__float_as_uint(d2);
                        if (myq >= 0 && j < cnt && vlab != ql && db <= qb && db > qin)
                            atomicMin(&best[myq],
                                      ((unsigned long long)__float_as_uint(d2) << 32) | __float_as_uint(v.w));
                    }
                    if (need) {
                        bk = best[lane];
                        bnd = min(key_bound(bk), lim);
                    }
                };
                // The gathering (a finite bound): every leaf whose box lies within
                // the largest bound of the open queries from their box, of another
                // component than all of them, and not inside all of their lower
                // bounds, by box-to-box tests of 64 nodes a step; then those
                // leaves nearest first, each tested per query.  A gathering that
                // outgrows the LDS lists leaves the pass to the per-query walk.
                bool done_by_gather = false;
                uint32_t B = open ? (uint32_t)bnd : 0u;
                for (int off = 32; off > 0; off >>= 1) B = max(B, (uint32_t)__shfl_xor((int)B, off));
                if (B < 0x7f800000u) {
                    float qlo[3] = {open ? px : INFINITY, open ? py : INFINITY, open ? pz : INFINITY};
                    float qhi[3] = {open ? px : -INFINITY, open ? py : -INFINITY, open ? pz : -INFINITY};
                    int32_t imin = open ? inner : 0x7fffffff;
                    for (int off = 32; off > 0; off >>= 1) {
    #pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            qlo[k] = fminf(qlo[k], __shfl_xor(qlo[k], off));
                            qhi[k] = fmaxf(qhi[k], __shfl_xor(qhi[k], off));
                        }
                        imin = min(imin, __shfl_xor(imin, off));
                    }
                    const int fo = (int)__builtin_ctzll(__ballot(open));
                    const int32_t lab0 = __builtin_amdgcn_readfirstlane(__shfl(mylab, fo));
                    const int32_t skip_lab = __ballot(open && mylab != lab0) == 0ull ? lab0 : (int32_t)0x80000000;
                    int head = 0, tail = gather_start(L, work), ngl = 0;
                    bool over = false;
                    __builtin_amdgcn_wave_barrier();
                    while (head < tail) {
                        const int take = min(64, tail - head);
                        const int nd = lane < take ? work[head + lane] : 0;
                        head += take;
                        st_nodes += (uint32_t)take;
                        bool pass_ = false, leaf = false;
                        uint32_t bd = 0;
                        if (lane < take) {
                            const Node x = tree[nd];
                            const int32_t xl = __float_as_int(x.pad[1]);
                            float m[3], u[3];
    #pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                m[k] = fmaxf(fmaxf(x.lo[k] - qhi[k], qlo[k] - x.hi[k]), 0.f);
                                u[k] = fmaxf(x.hi[k] - qlo[k], qhi[k] - x.lo[k]);
                            }
                            bd = __float_as_uint(D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]));
                            const float far = D == 3 ? d2_of(u[0], u[1], u[2]) : d2_of(u[0], u[1]);
                            leaf = nd >= L;
                            pass_ = xl != kNlabEmpty && xl != skip_lab && bd <= B &&
                                    (int32_t)__float_as_uint(far) > imin && !(leaf && nd == pre);
                        }
                        const uint64_t ml = __ballot(pass_ && leaf);
                        const int nlf = __popcll(ml);
                        int add;
                        if (!push_inner(pass_ && !leaf, nd, L, work, tail, kGatherWork, add) ||
                            ngl + nlf > kGatherLeaves) {
                            over = true;
                            break;
                        }
                        if (pass_ && leaf)
                            gathered[ngl + __popcll(ml & lanemask_lt())] = ((uint64_t)bd << 32) | (uint32_t)nd;
                        tail += add;
                        ngl += nlf;
                        __builtin_amdgcn_wave_barrier();
                    }
                    if (!over) {
                        done_by_gather = true;
                        auto try_leaf = [&](int node) {
                            const Node x = tree[node];
                            ++st_nodes;
                            const bool need = needed(x, __float_as_int(x.pad[1]), __float_as_uint(x.pad[0]), lim, bk);
                            const uint64_t needm = __ballot(need);
                            if (needm != 0ull) visit_leaf(node, need, needm);
                        };
                        if (pre >= 0) try_leaf(pre);
                        for (int c0 = 0; c0 < ngl; c0 += 64) {
                            const int cn = min(64, ngl - c0);
                            const uint64_t e = wave_sort_u64(lane < cn ? gathered[c0 + lane] : kEmptyKey);
                            for (int i = 0; i < cn; ++i) try_leaf(__builtin_amdgcn_readlane((int)(uint32_t)e, i));
                        }
                    }
                }
                // a gathering that outgrew its lists: the dispatch takes as long
                // as its slowest packet, and this packet's walk would visit the
                // union of its queries' neighbourhoods in one wave; the open
                // queries go to the per-query pass instead, one wave each, from
                // this pass on (no leaf has been visited in it yet, so their
                // bests are empty), while the round's list has room
                bool handed = false;        // this lane's query went to the per-query pass
                if (!done_by_gather && defer != nullptr && B < 0x7f800000u) {
                    const uint64_t entry = ((uint64_t)(uint32_t)pass << 32) | (uint32_t)p;
                    if (D == 3) {
                        handed = wave_append_capped_u64(open, entry, defer, &C->deferred, kDeferCap);
                    } else if (nq <= kDeferQueries) {
                        wave_append_u64(open, entry, defer, &C->deferred);
                        handed = open;
                    }
                    if (handed) {
                        open = false;
                        deferred = true;
                    }
                }
                const bool walk = !done_by_gather && __ballot(open) != 0ull;
                if (!done_by_gather && !walk) {
                    ++st_walks;
                } else if (walk) {
                    ++st_walks;
                    if (lane == 0) {
                        stk_id[0] = 1;
                        stk_ver[0] = -1;
                        if (pre >= 0) { stk_id[1] = pre; stk_ver[1] = -1; }
                    }
                    sp = pre >= 0 ? 2 : 1;
                    __builtin_amdgcn_wave_barrier();
                    while (sp > 0) {
                        --sp;
                        const int node = __builtin_amdgcn_readfirstlane(stk_id[sp]);
                        const int pver = __builtin_amdgcn_readfirstlane(stk_ver[sp]);
                        if (node == pre) {
                            if (pre_done) continue;
                            pre_done = true;
                        }
                        bool need;
                        if (pver == ver) {
                            need = (stk_mask[sp] >> lane) & 1ull;
                        } else {
                            {
                                const Node nd = tree[node];
                                need = needed(nd, __float_as_int(nd.pad[1]), __float_as_uint(nd.pad[0]), lim, bk);
                            }
                            ++st_nodes;
                        }
                        const uint64_t needm = __ballot(need);
                        if (needm == 0ull) continue;
                        if (node >= L) {
                            visit_leaf(node, need, needm);
                            ++ver;
                        } else {
                            // (rare: an unbounded pass, or a gathering that
                            // outgrew its lists) both children, the nearer last
                            // on the stack
                            const int first = 2 * node;
                            st_nodes += 2u;
                            const Node b0 = tree[first], b1 = tree[first + 1];
                            const uint64_t m0 =
                                __ballot(needed(b0, __float_as_int(b0.pad[1]), __float_as_uint(b0.pad[0]), lim, bk));
                            const uint64_t m1 =
                                __ballot(needed(b1, __float_as_int(b1.pad[1]), __float_as_uint(b1.pad[0]), lim, bk));
                            const float d0 = m0 ? cdist(b0) : -1.f, d1 = m1 ? cdist(b1) : -1.f;
                            // pushed farthest first, so the nearest pops first; at
                            // equal distance the lower node (smaller ids in a run
                            // of coincident points) pops first
                            const bool swap = d1 > d0;
                            const int ia = swap ? first + 1 : first, ib = swap ? first : first + 1;
                            const uint64_t ma = swap ? m1 : m0, mb = swap ? m0 : m1;
                            if (lane == 0) {
                                int t = sp;
                                if (ma) { stk_id[t] = ia; stk_ver[t] = ver; stk_mask[t] = ma; ++t; }
                                if (mb) { stk_id[t] = ib; stk_ver[t] = ver; stk_mask[t] = mb; ++t; }
                            }
                            sp += (ma != 0ull ? 1 : 0) + (mb != 0ull ? 1 : 0);
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                }
                // everything outside the component up to min(best, lim) has been
                // seen: best is exact when within lim; else nothing lies within
                // lim, and with lim = cbb the point stops on its lower bound
                bool fresh = false;
                uint64_t key1 = kEmptyKey;
                if (open) {
                    const uint64_t kb = best[lane];
                    if (key_bound(kb) <= lim) {
                        open = false;
                        if (kb != kEmptyKey) {
                            fresh = true;
                            key1 = (kb & 0xffffffff00000000ull) | (uint64_t)min(pid, (uint32_t)kb);
                        }
                    } else if (cbb <= cap) {
                        open = false;
                        bounded = true;
                        lbnd = cbb;
                    }
                }
                atomic_min_leader(comp1, mylab, key1, fresh);
                inner = max(inner, cap);
                r2 *= 8.f;
            }
            if (mine) left_open = open;
        }
        if (open_out != nullptr) {
            const uint64_t m = __ballot(left_open);
            if (qpp == 64) {
                if (lane == 0) open_out[pk] = m;
            } else if (lane == 0 && m != 0ull) {
                // several packets share a mask word (the compaction cleared it)
                atomicOr((unsigned long long*)&open_out[(pk * qpp) >> 6], (unsigned long long)(m << ((pk * qpp) & 63)));
            }
        }
        if (collect && lane == 0) {
            const int slot = (int)(pk % kWorkSlots);
            atomicAdd(&C->work[kWorkCompNodes][slot], (unsigned long long)st_nodes);
            atomicAdd(&C->work[kWorkCompLeaves][slot], (unsigned long long)st_leaves);
            atomicAdd(&C->work[kWorkCompPairs][slot], (unsigned long long)st_pairs);
            atomicAdd(&C->work[kWorkCompWalks][slot], (unsigned long long)st_walks);
        }
        if (active && !left_open && !deferred) {
            const uint64_t k = best[lane];
            if (bounded) {
                corr[p] = -1;
                corr_key[p] = ((uint64_t)(uint32_t)lbnd << 32) | 0xffffffffull;
            } else {
                corr[p] = k == kEmptyKey ? -2 : inv[(uint32_t)k];
                corr_key[p] = k;
            }
        }
    }
}

// The per-query pass of the component search: the queries a packet handed
// over (defer: position | pass << 32, count in C->deferred), one wave per
// query, from the pass it reached to the last.  The same search as
// k_comp_search's, for one query: the wave gathers the candidate leaves
// around the query point by box tests of 64 nodes a step, then visits them
// nearest first with one leaf point per lane and a wave minimum of the keys;
// a gathering that outgrows its lists walks the tree, nearest child first.
constexpr int kDeferWaves = 4;          // waves per block
constexpr int kDeferBlocks = 2048;      // a grid that fills the chip once; waves loop over the queries


__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = shfl_xor_u64(v, off);
        v = o < v ? o : v;
    }
    return v;
}

template <int D>
__global__ __launch_bounds__(64 * kDeferWaves) void k_comp_search_deferred(
    const float4* __restrict__ pts, const Node* __restrict__ nodes, const Cloud* __restrict__ clouds, int ncl,
    LeafTable lt, const uint64_t* __restrict__ defer, const int32_t* __restrict__ label,
    const int32_t* __restrict__ inv, float seed_scale, int32_t* __restrict__ corr,
    uint64_t* __restrict__ corr_key, uint64_t* __restrict__ comp1, const float* __restrict__ seed_r, Counters* C,
    int collect)
{
    if (C->done) return;
    // (in 3D the count may pass the list's cap: the lanes past it walked)
    const uint32_t cnt = D == 3 ? min(C->deferred, kDeferCap) : C->deferred;
    __shared__ int32_t work_w[kDeferWaves][kGatherWork];
    __shared__ unsigned long long gathered_w[kDeferWaves][kGatherLeaves];
    __shared__ int32_t stk_w[kDeferWaves][64];
    const int w = (int)threadIdx.x / 64;
    const int lane = __lane_id();
    int32_t* work = work_w[w];
    unsigned long long* gathered = gathered_w[w];
    int32_t* stk = stk_w[w];
    uint32_t st_nodes = 0, st_leaves = 0;
    uint64_t st_pairs = 0;
    for (uint32_t qi = blockIdx.x * kDeferWaves + w; qi < cnt; qi += gridDim.x * kDeferWaves) {
        const uint64_t e = defer[qi];
        const int64_t p = (int64_t)(uint32_t)e;
        const int pass0 = (int)(e >> 32);
        const float4 me = pts[p];
        const float px = me.x, py = me.y, pz = me.z;
        const uint32_t pid = __float_as_uint(me.w);
        const int32_t mylab = label[p];
        const uint64_t prev = corr_key[p];
        const int c = find_cloud(clouds, ncl, p);
        const int64_t cend = clouds[c].pos + clouds[c].n;
        const int64_t L = clouds[c].L, lfirst = clouds[c].lfirst, nl = clouds[c].nl;
        const Node* __restrict__ tree = nodes + 2 * clouds[c].lbase;
        // the seed and the state the earlier passes left, as in k_comp_search
        int32_t inner = -1;
        float r;
        {
            const int64_t lf = lt.of[p];
            const Node own = tree[L + (lf - lfirst)];
            const float cntl = (float)(lt.start[lf + 1] - lt.start[lf]);
            float vol = 1.f, ext = 0.f;
            for (int k = 0; k < D; ++k) {
                const float ee = own.hi[k] - own.lo[k];
                vol *= ee;
                ext = fmaxf(ext, ee);
            }
            float sp = vol > 0.f ? powf(vol / cntl, 1.f / D) : ext / cntl;
            if (!positive_finite(sp)) sp = 1e-6f;
            const bool has_lb = (uint32_t)(prev >> 32) != 0xffffffffu;
            const uint32_t dbits = (uint32_t)(prev >> 32);
            if (has_lb) inner = (int32_t)dbits - 1;
            r = seed_scale * sp;
            if (has_lb) r = fmaxf(r, 2.f * sqrtf(__uint_as_float(dbits)));
            if (seed_r != nullptr && C->round >= kLearnedFrom) {
                r = seed_r[p];
                if (has_lb) r = fmaxf(r, sqrtf(__uint_as_float(dbits)));
            }
            if (!positive_finite(r)) r = 1e-6f;
        }
        float r2 = r * r;
        for (int k = 0; k < pass0; ++k) {
            inner = max(inner, (int32_t)__float_as_uint(r2));
            r2 *= 8.f;
        }
        bool open = true, bounded = false;
        int32_t lbnd = 0;
        uint64_t bk = kEmptyKey;
        auto box_d2 = [&](const Node& x, float& far) {
            float m[3], u[3];
            const float lo[3] = {x.lo[0] - px, x.lo[1] - py, x.lo[2] - pz};
            const float hi[3] = {x.hi[0] - px, x.hi[1] - py, x.hi[2] - pz};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[k] = fmaxf(fmaxf(lo[k], -hi[k]), 0.f);
                u[k] = fmaxf(-lo[k], hi[k]);
            }
            far = D == 3 ? d2_of(u[0], u[1], u[2]) : d2_of(u[0], u[1]);
            return D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]);
        };
        for (int pass = pass0; pass <= kMaxCappedPasses && open; ++pass) {
            const bool last = pass == kMaxCappedPasses;
            // (other waves lower comp1[mylab] meanwhile, and the lanes of one
            // load can see different values: one lane's steers the wave)
            const int32_t cbb = key_bound(wave_first_u64(comp1[mylab]));
            const int32_t cap = last || cbb != 0x7fffffff ? 0x7fffffff : (int32_t)__float_as_uint(r2);
            if (cbb <= inner) {
                open = false;
                bounded = true;
                lbnd = inner;
                break;
            }
            const int32_t lim = min(cap, cbb);
            int32_t bnd = min(key_bound(bk), lim);
            // a node is needed when it holds another component, its nearest
            // point is within the limit and may beat the best so far under
            // (d2, id), and its farthest point lies beyond `inner`
            auto needed = [&](const Node& x) {
                const int32_t xl = __float_as_int(x.pad[1]);
                if (xl == mylab || xl == kNlabEmpty) return false;
                float far;
                const uint32_t lbb = __float_as_uint(box_d2(x, far));
                return (int32_t)lbb <= lim && (((uint64_t)lbb << 32) | __float_as_uint(x.pad[0])) < bk &&
                       (int32_t)__float_as_uint(far) > inner;
            };
            // one leaf point per lane, the wave's minimum key
            auto visit_leaf = [&](int node) {
                int64_t s0, s1;
                leaf_range(lt, lfirst, nl, cend, node - L, s0, s1);
                const int cn = (int)(s1 - s0);
                ++st_leaves;
                st_pairs += (uint64_t)cn;
                uint64_t key = kEmptyKey;
                if (lane < cn) {
                    const float4 v = pts[s0 + lane];
                    const float dx = v.x - px, dy = v.y - py, dz = v.z - pz;
                    const float d2 = D == 3 ? d2_of(dx, dy, dz) : d2_of(dx, dy);
                    const int32_t db = (int32_t)__float_as_uint(d2);
                    if (label[s0 + lane] != mylab && db <= bnd && db > inner)
                        key = ((uint64_t)__float_as_uint(d2) << 32) | __float_as_uint(v.w);
                }
                const uint64_t km = wave_min_u64(key);
                if (km < bk) {
                    bk = km;
                    bnd = min(key_bound(bk), lim);
                }
            };
            bool done_by_gather = false;
            if ((uint32_t)bnd < 0x7f800000u) {
                const uint32_t B = (uint32_t)bnd;
                int head = 0, tail = gather_start(L, work), ngl = 0;
                bool over = false;
                while (head < tail) {
                    const int take = min(64, tail - head);
                    const int nd = lane < take ? work[head + lane] : 0;
                    head += take;
                    st_nodes += (uint32_t)take;
                    bool pass_ = false, leaf = false;
                    uint32_t bd = 0;
                    if (lane < take) {
                        const Node x = tree[nd];
                        const int32_t xl = __float_as_int(x.pad[1]);
                        float far;
                        bd = __float_as_uint(box_d2(x, far));
                        leaf = nd >= L;
                        pass_ = xl != kNlabEmpty && xl != mylab && bd <= B && (int32_t)__float_as_uint(far) > inner;
                    }
                    const uint64_t ml = __ballot(pass_ && leaf);
                    const int nlf = __popcll(ml);
                    int add;
                    if (!push_inner(pass_ && !leaf, nd, L, work, tail, kGatherWork, add) ||
                        ngl + nlf > kGatherLeaves) {
                        over = true;
                        break;
                    }
                    if (pass_ && leaf)
                        gathered[ngl + __popcll(ml & lanemask_lt())] = ((uint64_t)bd << 32) | (uint32_t)nd;
                    tail += add;
                    ngl += nlf;
                    __builtin_amdgcn_wave_barrier();
                }
                if (!over) {
                    done_by_gather = true;
                    for (int c0 = 0; c0 < ngl; c0 += 64) {
                        const int cn = min(64, ngl - c0);
                        const uint64_t ge = wave_sort_u64(lane < cn ? gathered[c0 + lane] : kEmptyKey);
                        for (int i = 0; i < cn; ++i) {
                            const int node = __builtin_amdgcn_readlane((int)(uint32_t)ge, i);
                            const Node x = tree[node];
                            ++st_nodes;
                            if (needed(x)) visit_leaf(node);
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (!done_by_gather) {
                // the walk: a wave-uniform stack, nearest child popped first
                int sp = 1;
                if (lane == 0) stk[0] = 1;
                __builtin_amdgcn_wave_barrier();
                while (sp > 0) {
                    --sp;
                    const int node = __builtin_amdgcn_readfirstlane(stk[sp]);
                    __builtin_amdgcn_wave_barrier();
                    const Node x = tree[node];
                    ++st_nodes;
                    if (!needed(x)) continue;
                    if (node >= L) {
                        visit_leaf(node);
                        continue;
                    }
                    const int first = 2 * node;
                    const Node b0 = tree[first], b1 = tree[first + 1];
                    float f0, f1;
                    const float d0 = box_d2(b0, f0), d1 = box_d2(b1, f1);
                    // pushed farther first; at equal distance the lower node pops first
                    const bool swap = d1 >= d0;
                    if (lane == 0) {
                        stk[sp] = swap ? first + 1 : first;
                        stk[sp + 1] = swap ? first : first + 1;
                    }
                    sp += 2;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // as in k_comp_search: exact when the best is within lim, else
            // nothing lies within lim, and with lim = cbb the point stops on
            // its lower bound
            if (key_bound(bk) <= lim) {
                open = false;
                if (bk != kEmptyKey && lane == 0)
                    atomic_min_if_lower(&comp1[mylab],
                                        (bk & 0xffffffff00000000ull) | (uint64_t)min(pid, (uint32_t)bk));
            } else if (cbb <= cap) {
                open = false;
                bounded = true;
                lbnd = cbb;
            }
            inner = max(inner, cap);
            r2 *= 8.f;
        }
        if (lane == 0) {
            if (bounded) {
                corr[p] = -1;
                corr_key[p] = ((uint64_t)(uint32_t)lbnd << 32) | 0xffffffffull;
            } else {
                corr[p] = bk == kEmptyKey ? -2 : inv[(uint32_t)bk];
                corr_key[p] = bk;
            }
        }
    }
    if (collect && lane == 0) {
        const int slot = (blockIdx.x * kDeferWaves + w) % kWorkSlots;
        atomicAdd(&C->work[kWorkCompNodes][slot], (unsigned long long)st_nodes);
        atomicAdd(&C->work[kWorkCompLeaves][slot], (unsigned long long)st_leaves);
        atomicAdd(&C->work[kWorkCompPairs][slot], st_pairs);
    }
}

// Round 1 of the component search and its neighbour lists.  Every component
// is one point, so a point's nearest outside point is its nearest neighbour.
// One wave per 64 consecutive positions keeps each lane's K smallest keys
// (d2, id) in registers.  The leaves holding the packet's own points come
// first; then every lane's K-th key bounds its search.  The packet
// gathers, 64 nodes per step (one box-to-box test per lane), every other leaf
// whose box lies within the largest of those bounds of the packet's box, and
// visits them nearest first.  At a leaf, a lane takes part only while the
// leaf box may still beat its K-th key, and a leaf point enters the pair
// loop only when it lies within the bound of an 8-lane group of the packet
// that takes part (the group's box and its largest bound).  A packet whose
// gathering outgrows the LDS lists walks the hierarchy per lane instead,
// nearest child first.  In the pair loop each lane is one query and the
// leaf's points are read in turn at one address for the whole wave.
//
// The list holds keys only, biased by 2^52: (d2 bits << 32 | id) + 2^52 with
// d2 bits <= 0x7f800000 is a positive normal double whose order is the key's,
// so an insertion is a min/max network of K FP64 min/max pairs without
// branches or selects (the empty key is the largest finite double).  The
// positions come from inv at the end.
constexpr uint64_t kKnnBias = 1ull << 52;
constexpr uint64_t kKnnEmpty = 0x7fefffffffffffffull;
constexpr uint32_t kKnnBiasHi = (uint32_t)(kKnnBias >> 32);

// returns min(a, x) and leaves max(a, x) in x, for doubles that are never NaN:
// v_min_f64 / v_max_f64 without the canonicalising v_max_f64 the compiler
// puts before each input of fmin / fmax.  The minimum is written over a's
// register, so a network kd[i] = minmax_f64(kd[i], x) needs no copies.
__device__ __forceinline__ double minmax_f64(double a, double& x)
{
    double hi;
    asm("v_max_f64 %1, %0, %2\n\tv_min_f64 %0, %0, %2" : "+v"(a), "=&v"(hi) : "v"(x));
    x = hi;
    return a;
}
template <int D>
__global__ __launch_bounds__(64) void k_knn_first(int64_t n, const float4* __restrict__ pts,
                                                 const Node* __restrict__ nodes, const Cloud* __restrict__ clouds,
                                                 int ncl, int bucket, LeafTable lt, const int32_t* __restrict__ inv,
                                                 int32_t* __restrict__ corr, uint64_t* __restrict__ corr_key,
                                                 NbrList nb, Counters* C, int collect)
{
    __shared__ int32_t stk_id[64];
    __shared__ int32_t stk_ver[64];
    __shared__ unsigned long long stk_mask[64];
    __shared__ int32_t work[kKnnGatherWork];
    __shared__ unsigned long long gathered[kKnnGatherLeaves];
    __shared__ float4 gbox[16];             // per 8-lane group: lo (xyz), hi (xyz)
    if (C->done) return;                    // a build whose input was refused on the device
    const int lane = __lane_id();
    if (blockIdx.x == 0 && lane == 0) C->knn_round1 = 1u;
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;
    const bool active = p < n;
    const float4 me = pts[active ? p : 0];
    const float px = me.x, py = me.y, pz = me.z;
    const int my_cl = active ? find_cloud(clouds, ncl, p) : -1;
    constexpr int K = knn_len(D);   // list length
    double kd[K];                   // biased keys as doubles, ascending
#pragma unroll
    for (int i = 0; i < K; ++i) kd[i] = __longlong_as_double((long long)kKnnEmpty);
    auto last = [&]() { return (uint64_t)__double_as_longlong(kd[K - 1]); };
    // sorted insertion of a biased key (keys are distinct: the id is in the
    // low word)
    auto insert = [&](uint64_t key) {
        double x = __longlong_as_double((long long)key);
#pragma unroll
        for (int i = 0; i < K; ++i) kd[i] = minmax_f64(kd[i], x);
    };
    // the K-th d2 bits (an empty list: above every finite d2)
    auto bound = [&]() { return (uint32_t)(last() >> 32) - kKnnBiasHi; };
    uint32_t st_nodes = 0, st_leaves = 0, st_steps = 0, st_walks = 0;   // work counters of this packet
    uint64_t pending = __ballot(active);
    while (pending) {
        const int cl = __builtin_amdgcn_readfirstlane(__shfl(my_cl, (int)__builtin_ctzll(pending)));
        const uint64_t minem = __ballot(my_cl == cl);
        const bool mine = my_cl == cl;
        pending &= ~minem;
        const int64_t cend = clouds[cl].pos + clouds[cl].n;
        const int64_t L = clouds[cl].L, lfirst = clouds[cl].lfirst, nl = clouds[cl].nl;
        const Node* __restrict__ tree = nodes + 2 * clouds[cl].lbase;
        // the box of the group's points, and of each 8-lane group's
        float qlo[3] = {mine ? px : INFINITY, mine ? py : INFINITY, mine ? pz : INFINITY};
        float qhi[3] = {mine ? px : -INFINITY, mine ? py : -INFINITY, mine ? pz : -INFINITY};
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                qlo[k] = fminf(qlo[k], __shfl_xor(qlo[k], off));
                qhi[k] = fmaxf(qhi[k], __shfl_xor(qhi[k], off));
            }
            if (off == 4 && (lane & 7) == 0) {
                gbox[2 * (lane >> 3)] = make_float4(qlo[0], qlo[1], qlo[2], 0.f);
                gbox[2 * (lane >> 3) + 1] = make_float4(qhi[0], qhi[1], qhi[2], 0.f);
            }
        }
        __builtin_amdgcn_wave_barrier();
        auto needed = [&](const Node& nd, uint32_t nm) {
            if (!mine) return false;
            const float q[3] = {px, py, pz};
            float m[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) m[k] = fmaxf(fmaxf(nd.lo[k] - q[k], q[k] - nd.hi[k]), 0.f);
            const float lb = D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]);
            return (((uint64_t)(__float_as_uint(lb) + kKnnBiasHi) << 32) | nm) < last();
        };
        // one leaf point per candidate step, at one address for the wave
        auto pairs = [&](int64_t s0, uint64_t cand, bool own) {
            ++st_leaves;
            st_steps += (uint32_t)__popcll(cand);
            while (cand) {
                const int j = (int)__builtin_ctzll(cand);
                cand &= cand - 1;
                const float4 v = pts[s0 + j];
                const float dx = v.x - px, dy = v.y - py, dz = v.z - pz;
                const float d2 = D == 3 ? d2_of(dx, dy, dz) : d2_of(dx, dy);
                const uint32_t db = __float_as_uint(d2);
                const uint64_t key = ((uint64_t)(db + kKnnBiasHi) << 32) | __float_as_uint(v.w);
                if (mine && key < last() && (!own || s0 + j != p)) insert(key);
            }
        };
        // the packet's own leaves: a leaf point farther from the packet's box
        // than every lane's K-th d2 cannot enter a list
        auto scan_own = [&](int node) {
            int64_t s0, s1;
            leaf_range(lt, lfirst, nl, cend, node - L, s0, s1);
            const int cnt = (int)(s1 - s0);
            uint32_t t = mine ? bound() : 0u;
            for (int off = 32; off > 0; off >>= 1) t = max(t, (uint32_t)__shfl_xor((int)t, off));
            bool keep = false;
            if (lane < cnt) {
                const float4 w = pts[s0 + lane];
                const float wq[3] = {w.x, w.y, w.z};
                float m[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) m[k] = fmaxf(fmaxf(qlo[k] - wq[k], wq[k] - qhi[k]), 0.f);
                const float lb = D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]);
                keep = __float_as_uint(lb) <= t;
            }
            pairs(s0, __ballot(keep), true);
        };
        // another leaf, for the lanes of needm: a leaf point is paired when it
        // lies within the largest K-th d2 of a taking-part group's lanes
        // from that group's box
        auto scan_other = [&](int node, bool need, uint64_t needm) {
            int64_t s0, s1;
            leaf_range(lt, lfirst, nl, cend, node - L, s0, s1);
            const int cnt = (int)(s1 - s0);
            const uint32_t t = group8_max_u32(need ? bound() : 0u);
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane < cnt) w = pts[s0 + lane];
            bool keep = false;
#pragma unroll 1
            for (int g = 0; g < 8; ++g) {
                if (((needm >> (8 * g)) & 0xffull) == 0ull) continue;
                const uint32_t tg = (uint32_t)__builtin_amdgcn_readlane((int)t, 8 * g);
                const float4 lo = gbox[2 * g], hi = gbox[2 * g + 1];
                const float m0 = fmaxf(fmaxf(lo.x - w.x, w.x - hi.x), 0.f);
                const float m1 = fmaxf(fmaxf(lo.y - w.y, w.y - hi.y), 0.f);
                const float m2 = fmaxf(fmaxf(lo.z - w.z, w.z - hi.z), 0.f);
                const float lb = D == 3 ? d2_of(m0, m1, m2) : d2_of(m0, m1);
                keep = keep || __float_as_uint(lb) <= tg;
            }
            pairs(s0, __ballot(keep && lane < cnt), false);
        };
        // the leaves holding the packet's points of this cloud
        const int first = (int)__builtin_ctzll(minem), last = 63 - (int)__builtin_clzll(minem);
        const int64_t pf = (int64_t)blockIdx.x * 64 + first, pl = (int64_t)blockIdx.x * 64 + last;
        const int leaf_lo = (int)(L + (lt.of[pf] - lfirst)), leaf_hi = (int)(L + (lt.of[pl] - lfirst));
        for (int lf = leaf_lo; lf <= leaf_hi; ++lf) scan_own(lf);
        // gather the other leaves within the largest K-th d2 of the packet's box
        uint32_t R = mine ? bound() : 0u;
        for (int off = 32; off > 0; off >>= 1) R = max(R, (uint32_t)__shfl_xor((int)R, off));
        int head = 0, tail = gather_start(L, work), ngl = 0;
        bool over = false;
        __builtin_amdgcn_wave_barrier();
        while (head < tail) {
            const int take = min(64, tail - head);
            const int nd = lane < take ? work[head + lane] : 0;
            head += take;
            st_nodes += (uint32_t)take;
            bool pass = false, leaf = false;
            uint32_t bd = 0;
            if (lane < take) {
                const Node x = tree[nd];
                float m[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) m[k] = fmaxf(fmaxf(x.lo[k] - qhi[k], qlo[k] - x.hi[k]), 0.f);
                bd = __float_as_uint(D == 3 ? d2_of(m[0], m[1], m[2]) : d2_of(m[0], m[1]));
                leaf = nd >= L;
                pass = bd <= R && !(leaf && nd >= leaf_lo && nd <= leaf_hi);
            }
            const uint64_t ml = __ballot(pass && leaf);
            const int nl = __popcll(ml);
            int add;
            if (!push_inner(pass && !leaf, nd, L, work, tail, kKnnGatherWork, add) || ngl + nl > kKnnGatherLeaves) {
                over = true;
                break;
            }
            if (pass && leaf) gathered[ngl + __popcll(ml & lanemask_lt())] = ((uint64_t)bd << 32) | (uint32_t)nd;
            tail += add;
            ngl += nl;
            __builtin_amdgcn_wave_barrier();
        }
        if (!over) {
            // nearest first, 64 leaves at a time
            for (int c0 = 0; c0 < ngl; c0 += 64) {
                const int cn = min(64, ngl - c0);
                const uint64_t e = wave_sort_u64(lane < cn ? gathered[c0 + lane] : kEmptyKey);
                for (int i = 0; i < cn; ++i) {
                    const int node = __builtin_amdgcn_readlane((int)(uint32_t)e, i);
                    const Node x = tree[node];
                    ++st_nodes;
                    const bool need = needed(x, __float_as_uint(x.pad[0]));
                    const uint64_t needm = __ballot(need);
                    if (needm != 0ull) scan_other(node, need, needm);
                }
            }
            continue;
        }
        // the per-lane walk: from the root, nearest child first (from the
        // packet's centroid), the own leaves skipped
        ++st_walks;
        float cx = mine ? px : 0.f, cy = mine ? py : 0.f, cz = mine ? pz : 0.f;
        float na = mine ? 1.f : 0.f;
        for (int off = 32; off > 0; off >>= 1) {
            cx += __shfl_xor(cx, off);
            cy += __shfl_xor(cy, off);
            cz += __shfl_xor(cz, off);
            na += __shfl_xor(na, off);
        }
        cx /= na; cy /= na; cz /= na;
        auto cdist = [&](const Node& x) {
            const float dxx = fmaxf(fmaxf(x.lo[0] - cx, cx - x.hi[0]), 0.f);
            const float dyy = fmaxf(fmaxf(x.lo[1] - cy, cy - x.hi[1]), 0.f);
            const float dzz = fmaxf(fmaxf(x.lo[2] - cz, cz - x.hi[2]), 0.f);
            return dxx * dxx + dyy * dyy + dzz * dzz;
        };
        int sp = 1;
        int ver = 0;
        if (lane == 0) {
            stk_id[0] = 1;
            stk_ver[0] = -1;
        }
        __builtin_amdgcn_wave_barrier();
        while (sp > 0) {
            --sp;
            const int node = __builtin_amdgcn_readfirstlane(stk_id[sp]);
            const int pver = __builtin_amdgcn_readfirstlane(stk_ver[sp]);
            bool need;
            if (pver == ver) {
                need = (stk_mask[sp] >> lane) & 1ull;
            } else {
                const Node nd = tree[node];
                need = needed(nd, __float_as_uint(nd.pad[0]));
                ++st_nodes;
            }
            const uint64_t needm = __ballot(need);
            if (needm == 0ull) continue;
            if (node >= L) {
                if (node >= leaf_lo && node <= leaf_hi) continue;
                scan_other(node, need, needm);
                ++ver;
            } else {
                const bool leafkids = 2 * (int64_t)node >= L;
                const int fc = leafkids ? 2 * node : 4 * node;
                const int nk = leafkids ? 2 : 4;
                st_nodes += (uint32_t)nk;
                uint64_t mk[4];
                float dk[4];
                int id[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    mk[i] = 0ull;
                    dk[i] = -1.f;
                    id[i] = fc + i;
                    if (i < nk) {
                        const Node bx = tree[fc + i];
                        mk[i] = __ballot(needed(bx, __float_as_uint(bx.pad[0])));
                        if (mk[i]) dk[i] = cdist(bx);
                    }
                }
                // pushed farthest first, so the nearest pops first
                auto ce = [&](int a, int b) {
                    if (dk[a] < dk[b] || (dk[a] == dk[b] && id[a] < id[b])) {
                        const float td = dk[a]; dk[a] = dk[b]; dk[b] = td;
                        const uint64_t tm = mk[a]; mk[a] = mk[b]; mk[b] = tm;
                        const int ti = id[a]; id[a] = id[b]; id[b] = ti;
                    }
                };
                ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2);
                if (lane == 0) {
                    int t = sp;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (mk[i]) { stk_id[t] = id[i]; stk_ver[t] = ver; stk_mask[t] = mk[i]; ++t; }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) sp += mk[i] != 0ull ? 1 : 0;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (collect && lane == 0) {
        const int slot = blockIdx.x % kWorkSlots;
        atomicAdd(&C->work[kWorkKnnNodes][slot], (unsigned long long)st_nodes);
        atomicAdd(&C->work[kWorkKnnLeaves][slot], (unsigned long long)st_leaves);
        atomicAdd(&C->work[kWorkKnnPairs][slot], 64ull * st_steps);
        atomicAdd(&C->work[kWorkKnnWalks][slot], (unsigned long long)st_walks);
    }
    if (!active) return;
    int32_t pos0 = -2;
    uint64_t key0 = kEmptyKey;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const uint64_t b = (uint64_t)__double_as_longlong(kd[i]);
        const bool has = b != kKnnEmpty;
        const uint64_t key = has ? b - kKnnBias : kEmptyKey;
        const int32_t pos = has ? inv[(uint32_t)key] : -1;
        if (i == 0 && has) {
            pos0 = pos;
            key0 = key;
        }
        nb.pos[(int64_t)i * n + p] = pos;
        nb.key[(int64_t)i * n + p] = key;
    }
    corr[p] = pos0;
    corr_key[p] = key0;
    nb.cur[p] = 0;
}

// ---- cycle rule of the 2.5D mode (CycleState).  Roots are the points whose
// flattened label is their own position.  In round 1 every component is one
// point and its chosen edge is its nearest neighbour, whose choices close
// only mutual pairs (the edges along a cycle of choices could not increase),
// so the rule starts at round 2.
//
// k_cycle_init: per root, its target j0 (itself when it has no outgoing
// candidate or is in a mutual pair), both doubling buffers, onc = cut = 0;
// the roots with another target are listed (cy.roots, C->nroots), one
// block-wide reservation per chunk of kCyChunk points.
constexpr int kCyChunk = 8192;
constexpr int kCyPer = kCyChunk / kBlock;      // points per thread, one bit each

__global__ __launch_bounds__(kBlock) void k_cycle_init(int64_t n, const int32_t* __restrict__ label,
                                                      const uint64_t* __restrict__ comp1,
                                                      const uint32_t* __restrict__ comp2, CycleState cy,
                                                      Counters* C)
{
    if (C->done || C->round <= 1u) return;
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t base;
    const int64_t c0 = (int64_t)blockIdx.x * kCyChunk;
    uint32_t bits = 0;
    for (int i = 0; i < kCyPer; ++i) {
        const int64_t r = c0 + (int64_t)i * kBlock + threadIdx.x;
        if (r >= n || label[r] != (int32_t)r) continue;
        int32_t t = (int32_t)r;
        const uint64_t k1 = comp1[r];
        const uint32_t k2 = comp2[r];
        if (k1 != kEmptyKey) {
            // the chosen edge (d2, min id, max id): its end outside the component
            const int32_t la = label[cy.inv[(uint32_t)k1]], lb = label[cy.inv[k2]];
            const int32_t o = la == (int32_t)r ? lb : la;
            if (!(comp1[o] == k1 && comp2[o] == k2)) t = o;     // else a mutual pair
        }
        const uint64_t h = t == (int32_t)r ? 0ull : k1;
        const uint32_t l = t == (int32_t)r ? 0u : k2;
        cy.j0[r] = t;
        cy.j[0][r] = t;
        cy.j[1][r] = t;
        cy.kh[0][r] = h;
        cy.kh[1][r] = h;
        cy.kl[0][r] = l;
        cy.kl[1][r] = l;
        cy.onc[r] = 0;
        cy.cut[r] = 0;
        if (t != (int32_t)r) bits |= 1u << i;
    }
    // block-exclusive offsets of the listed roots, in thread order
    const uint32_t cnt = (uint32_t)__popc(bits);
    uint32_t incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)incl, off);
        if (__lane_id() >= off) incl += o;
    }
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            const uint32_t x = wsum[k];
            wsum[k] = tot;
            tot += x;
        }
        base = tot ? atomicAdd(&C->nroots, tot) : 0u;
    }
    __syncthreads();
    uint32_t at = base + wsum[w] + incl - cnt;
    for (int i = 0; i < kCyPer; ++i)
        if ((bits >> i) & 1u) cy.roots[at++] = (int32_t)(c0 + (int64_t)i * kBlock + threadIdx.x);
}

// number of doubling steps for R listed roots: 2^S > R covers any path of
// listed roots, which ends on a cycle or at a root with no target
__device__ __forceinline__ int cycle_steps(uint32_t R, int max_steps)
{
    int S = 0;
    while (S < max_steps && (1ull << S) <= (uint64_t)R) ++S;
    return S;
}

// doubling step k (reads buffer k & 1): hop 2^k further, keep the largest
// edge key on the way; steps past cycle_steps do nothing
__global__ __launch_bounds__(kBlock) void k_cycle_step(CycleState cy, int k, const Counters* C)
{
    if (C->done || C->round <= 1u) return;
    const uint32_t R = C->nroots;
    if (k >= cycle_steps(R, cy.steps)) return;
    const int src = k & 1, dst = src ^ 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)R;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = cy.roots[t];
        const int32_t j = cy.j[src][r];
        uint64_t h = cy.kh[src][r];
        uint32_t l = cy.kl[src][r];
        const uint64_t h2 = cy.kh[src][j];
        const uint32_t l2 = cy.kl[src][j];
        if (h2 > h || (h2 == h && l2 > l)) {
            h = h2;
            l = l2;
        }
        cy.j[dst][r] = cy.j[src][j];
        cy.kh[dst][r] = h;
        cy.kl[dst][r] = l;
    }
}

// after the steps, every listed root's hop target lies on a cycle (or is a
// root with no target), and every root of a cycle is such a target
__global__ __launch_bounds__(kBlock) void k_cycle_mark(CycleState cy, const Counters* C)
{
    if (C->done || C->round <= 1u) return;
    const uint32_t R = C->nroots;
    const int fin = cycle_steps(R, cy.steps) & 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)R;
         t += (int64_t)gridDim.x * blockDim.x)
        cy.onc[cy.j[fin][cy.roots[t]]] = 1;
}

__global__ __launch_bounds__(kBlock) void k_cycle_cut(const uint64_t* __restrict__ comp1,
                                                     const uint32_t* __restrict__ comp2, CycleState cy,
                                                     Counters* C)
{
    if (C->done || C->round <= 1u) return;
    const uint32_t R = C->nroots;
    const int fin = cycle_steps(R, cy.steps) & 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)R;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = cy.roots[t];
        const bool cut = cy.onc[r] && comp1[r] == cy.kh[fin][r] && comp2[r] == cy.kl[fin][r];
        cy.cut[r] = cut ? 1 : 0;
        if (cut) atomicAdd(&C->cycles_cut, 1u);
    }
}

// flatten (EMSTOperators.h:412-433): chase label2 to the root, with path
// halving on label2 (every write stores an ancestor, so concurrent chases stay
// valid and label2 stays a parent forest); reset the component minima for the
// next round.  Only this round's roots (label[p] == p) can hold minima (every
// update goes to a point's root), and a label that stays is not rewritten:
// later rounds absorb few roots, and the stores of unchanged values took most
// of the pass.
// Each launch takes at most kFlattenSteps steps per point; a point still off
// its root raises C->unflat and the host launches the kernel again, so every
// wave ends even if the hook forest were malformed.
constexpr int kFlattenSteps = 4096;

__global__ __launch_bounds__(kBlock) void k_round_flatten(int64_t n, int32_t* __restrict__ label,
                                                         int32_t* label2,
                                                         uint64_t* __restrict__ comp1,
                                                         uint32_t* __restrict__ comp2, Counters* C)
{
    if (C->done == 1u) return;      // a paused round (done == 2) still flattens
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t was = label[p];
    int32_t r = label2[p];
    int32_t nx = label2[r];
    int steps = 0;
    while (nx != r && steps < kFlattenSteps) {
        const int32_t nn = label2[nx];
        label2[r] = nn;
        r = nn;
        nx = label2[r];
        ++steps;
    }
    if (nx != r) atomicOr(&C->unflat, 1u);
    // p itself now points at the root (or the farthest ancestor reached), so
    // the next round's chase from p starts there
    if (steps > 0) label2[p] = r;
    if (r != was) label[p] = r;
    if (was == (int32_t)p) {
        comp1[p] = kEmptyKey;
        comp2[p] = 0xffffffffu;
    }
}

// ------------------------------------------------------------- weight
template <int D>
__global__ __launch_bounds__(kBlock) void k_weight(const float* __restrict__ xyz,
                                                  const int32_t* __restrict__ eu,
                                                  const int32_t* __restrict__ ev, int64_t m,
                                                  double* __restrict__ partial)
{
    __shared__ double sh[kBlock / 64];
    double s = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        if (eu[i] < 0) continue;
        const float* a = xyz + (int64_t)eu[i] * D;
        const float* b = xyz + (int64_t)ev[i] * D;
        const float dx = a[0] - b[0], dy = a[1] - b[1];
        const float d2 = D == 3 ? d2_of(dx, dy, a[2] - b[2]) : d2_of(dx, dy);
        s += sqrt((double)d2);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
        partial[blockIdx.x] = t;
    }
}

// Per cloud (one block each): FP64 weight of its edges in input coordinates,
// edge count, last round that added an edge.  Edges sit at the sorted
// position of the root they absorbed, inside the cloud's range.
template <int D>
__global__ __launch_bounds__(kBlock) void k_cloud_results(const float* __restrict__ xyz,
                                                         const Cloud* __restrict__ cl,
                                                         const int32_t* __restrict__ eu,
                                                         const int32_t* __restrict__ ev,
                                                         const int32_t* __restrict__ eround,
                                                         double* __restrict__ weight,
                                                         int32_t* __restrict__ edges,
                                                         int32_t* __restrict__ last_round)
{
    __shared__ double shw[kBlock / 64];
    __shared__ int32_t shc[kBlock / 64], shr[kBlock / 64];
    const int c = blockIdx.x;
    const int64_t a = cl[c].pos, b = a + cl[c].n;
    double s = 0;
    int32_t cnt = 0, lr = 0;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
        if (eu[i] < 0) continue;
        const float* p = xyz + (int64_t)eu[i] * D;
        const float* q = xyz + (int64_t)ev[i] * D;
        const float dx = p[0] - q[0], dy = p[1] - q[1];
        const float d2 = D == 3 ? d2_of(dx, dy, p[2] - q[2]) : d2_of(dx, dy);
        s += sqrt((double)d2);
        ++cnt;
        lr = max(lr, eround[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        cnt += __shfl_xor(cnt, off);
        lr = max(lr, __shfl_xor(lr, off));
    }
    if (__lane_id() == 0) {
        shw[threadIdx.x / 64] = s;
        shc[threadIdx.x / 64] = cnt;
        shr[threadIdx.x / 64] = lr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        int32_t tc = 0, tr = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            t += shw[w];
            tc += shc[w];
            tr = max(tr, shr[w]);
        }
        weight[c] = t;
        edges[c] = tc;
        last_round[c] = tr;
    }
}

// ------------------------------------------------------------ outputs
__global__ __launch_bounds__(kBlock) void k_labels_out(int64_t n, const int32_t* __restrict__ label,
                                                      const float4* __restrict__ pts,
                                                      const int32_t* __restrict__ inv,
                                                      int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (int32_t)__float_as_uint(pts[label[inv[i]]].w);
}

__global__ __launch_bounds__(kBlock) void k_cands_out(int K, int64_t n, const uint64_t* __restrict__ ck,
                                                     const int32_t* __restrict__ inv,
                                                     int32_t* __restrict__ cand, float* __restrict__ cd2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t s = inv[i];
    for (int k = 0; k < K; ++k) {
        const uint64_t key = ck[(int64_t)k * n + s];
        cand[i * K + k] = key == kEmptyKey ? -1 : (int32_t)(uint32_t)key;
        cd2[i * K + k] = key == kEmptyKey ? INFINITY : __uint_as_float((uint32_t)(key >> 32));
    }
}

__global__ __launch_bounds__(kBlock) void k_edge_keys(const int32_t* __restrict__ eu,
                                                     const int32_t* __restrict__ ev, int64_t m,
                                                     uint64_t* __restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    keys[i] = eu[i] < 0 ? ~0ull : (((uint64_t)(uint32_t)eu[i] << 32) | (uint32_t)ev[i]);
}

__global__ __launch_bounds__(kBlock) void k_edges_out(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ idx,
                                                     const float* __restrict__ ed2, int64_t m,
                                                     int32_t* __restrict__ u, int32_t* __restrict__ v,
                                                     float* __restrict__ d2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    u[i] = (int32_t)(k >> 32);
    v[i] = (int32_t)(uint32_t)k;
    d2[i] = ed2[idx[i]];
}

// ------------------------------------------------ reference cell grid
// findCell (ViewGrid.h:339-359): p1 = FEuclid(FDual(0)), size_slab = 2R;
// li = (float)(p - p1) + R in double; cell = floor(li / 2R).
__global__ __launch_bounds__(kBlock) void k_cell_grid(const float* __restrict__ resc, int64_t n, int dim,
                                                     int pstride, const int32_t* __restrict__ geo,
                                                     int32_t* __restrict__ cell_of,
                                                     uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ cell_count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int R = geo[9];
    bool ok = true;
    int32_t c[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) {
        const int32_t PC = geo[3 + k], pc = geo[6 + k];
        const float p1 = (float)((PC % 2) * R - PC * R + pc);
        const float p2 = (float)((2 + PC % 2) * R - PC * R + pc);
        const double size_slab = (double)(p2 - p1);
        const float dif = resc[i * pstride + k] - p1;
        const double li = (double)dif + size_slab / 2;
        c[k] = (int32_t)floor(li / size_slab);
        ok = ok && c[k] >= 0 && c[k] < geo[k];
    }
    for (int k = 0; k < 3; ++k) cell_of[i * 3 + k] = ok ? c[k] : -1;
    const int64_t ncell = (int64_t)geo[0] * geo[1] * geo[2];
    const int64_t lin = ok ? (int64_t)c[0] + (int64_t)geo[0] * ((int64_t)c[1] + (int64_t)geo[1] * c[2]) : ncell;
    keys[i] = (uint64_t)lin;
    if (ok) atomicAdd(&cell_count[lin], 1u);
}

inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ================================================================ launchers
void launch_bbox(const float* xyz, const Cloud* cl, int ncl, int64_t N, int dim, Params* P, hipStream_t s)
{
    hipLaunchKernelGGL(k_bbox_init, dim3(nblocks(ncl)), dim3(kBlock), 0, s, P, ncl);
    if (N <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>(nblocks(N), 2048);
    if (dim == 3) hipLaunchKernelGGL(k_bbox<3>, dim3(g), dim3(kBlock), 0, s, xyz, cl, ncl, N, P);
    else hipLaunchKernelGGL(k_bbox<2>, dim3(g), dim3(kBlock), 0, s, xyz, cl, ncl, N, P);
}

void launch_params(Params* P, const Cloud* cl, int ncl, int dim, int mbits, hipStream_t s)
{
    hipLaunchKernelGGL(k_params, dim3(nblocks(ncl)), dim3(kBlock), 0, s, P, cl, ncl, dim, mbits);
}

void launch_rescale_morton(const float* xyz, const Cloud* cl, int ncl, int64_t N, int dim, int mbits,
                           const Params* P, float* resc, uint64_t* keys, hipStream_t s)
{
    if (N <= 0) return;
    if (dim == 3)
        hipLaunchKernelGGL(k_rescale_morton<3>, dim3(nblocks(N)), dim3(kBlock), 0, s, xyz, cl, ncl, N, mbits, P,
                           resc, keys);
    else
        hipLaunchKernelGGL(k_rescale_morton<2>, dim3(nblocks(N)), dim3(kBlock), 0, s, xyz, cl, ncl, N, mbits, P,
                           resc, keys);
}

size_t scan_tmp_words(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, hipStream_t s)
{
    if (n <= 0) return;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kBlock), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kBlock), 0, s, out, n, tmp);
}

size_t radix_hist_words(int64_t n)
{
    const int64_t nblk = (n + kRsTile - 1) / kRsTile;
    return (size_t)(256 * std::max<int64_t>(nblk, 1));
}

void radix_sort(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int bits,
                bool vals_identity, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s)
{
    if (n <= 0) return;
    const int64_t nblk = (n + kRsTile - 1) / kRsTile;
    uint64_t* ka = k0; uint32_t* va = v0;
    uint64_t* kb = k1; uint32_t* vb = v1;
    bool identity = vals_identity;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8, ++passes) {
        hipLaunchKernelGGL(k_rs_count, dim3((unsigned)nblk), dim3(kBlock), 0, s, ka, n, shift, hist, nblk);
        exclusive_scan_u32(hist, hist, 256 * nblk, scan_tmp, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nblk), dim3(kBlock), 0, s, ka, va, kb, vb, n,
                           shift, hist, nblk, identity ? 1 : 0);
        identity = false;
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != k0) {
        hipMemcpyAsync(k0, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(v0, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    } else if (passes == 0 && vals_identity) {
        // no pass ran: values are the identity permutation
        hipMemcpyAsync(v0, v0, 0, hipMemcpyDeviceToDevice, s);
    }
}

void launch_gather_sorted(const float* resc, const uint32_t* perm, int64_t n, int dim, float4* pts,
                          int32_t* inv, uint32_t* ids, hipStream_t s)
{
    if (n <= 0) return;
    if (dim == 3)
        hipLaunchKernelGGL(k_gather_sorted<3>, dim3(nblocks(n)), dim3(kBlock), 0, s, resc, perm, n, pts, inv, ids);
    else
        hipLaunchKernelGGL(k_gather_sorted<2>, dim3(nblocks(n)), dim3(kBlock), 0, s, resc, perm, n, pts, inv, ids);
}

void launch_dup_groups(const uint64_t* keys, const float4* pts, int64_t n, int32_t* rep,
                       int32_t* runs, int32_t* list, int32_t* gslot, Counters* C, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dup_runs, dim3(nblocks(n)), dim3(kBlock), 0, s, keys, n, rep, runs, C);
    const unsigned g = (unsigned)std::min<int64_t>(nblocks(n), 1024);
    hipLaunchKernelGGL(k_dup_group, dim3(g), dim3(kBlock), 0, s, keys, pts, n, rep, runs, list, gslot, C);
    hipLaunchKernelGGL(k_dup_big, dim3((unsigned)std::min<int64_t>(nblocks(n), 512)), dim3(kBlock), 0, s, pts, n,
                       rep, runs, list, C);
    hipLaunchKernelGGL(k_dup_finish, dim3(1), dim3(1), 0, s, C);
}

void launch_build_tree(const float4* pts, const Cloud* cl, int ncl, int64_t nleaves, int64_t Lmax, LeafTable lt,
                       Node* nodes, const int32_t* rep, hipStream_t s)
{
    hipLaunchKernelGGL(k_leaves, dim3(nblocks(nleaves)), dim3(kBlock), 0, s, pts, cl, ncl, nleaves, lt, nodes,
                       rep);
    for (int64_t lo = Lmax / 2; lo >= 1; lo /= 2)
        hipLaunchKernelGGL(k_level, dim3(nblocks((int64_t)ncl * lo)), dim3(kBlock), 0, s, nodes, cl, ncl, lo);
}

void launch_cone_search(int dim, int cdim, const float4* pts, const Node* nodes, const Cloud* cl, int ncl,
                        int bucket, LeafTable lt, int64_t n, const int32_t* qlist, int64_t nq, const int32_t* label,
                        bool check_labels, const int32_t* inv, const float* seed_r, float seed_scale,
                        uint64_t* cand_key, int32_t* cand_pos, int32_t* revive, int32_t* first_corr,
                        uint64_t* first_key, Counters* C, bool collect_stats, DupState dup,
                        const SplitScratch* split, hipStream_t s)
{
    if (nq <= 0) return;
    const unsigned g = nblocks(nq, 64);
    const int cs = collect_stats ? 1 : 0;
    auto kern = dim == 3 ? (cdim == 2 ? k_cone_search<3, 2> : k_cone_search<3, 3>) : k_cone_search<2, 2>;
    if (!split || qlist || split->after < 0 || split->after >= kMaxCappedPasses) {
        hipLaunchKernelGGL(kern, dim3(g), dim3(64), 0, s, pts, nodes, cl, ncl, bucket, lt, n, qlist, nq, label,
                           check_labels ? 1 : 0, inv, seed_r, seed_scale, cand_key, cand_pos, revive, first_corr,
                           first_key, C, cs, dup, 0, kMaxCappedPasses, (uint64_t*)nullptr,
                           (const uint32_t*)nullptr, (const uint64_t*)nullptr, (int64_t)0);
        return;
    }
    // passes 0..after over every query, then the open ones compacted into full waves
    const int64_t P = (int64_t)g;
    hipLaunchKernelGGL(kern, dim3(g), dim3(64), 0, s, pts, nodes, cl, ncl, bucket, lt, n, qlist, nq, label,
                       check_labels ? 1 : 0, inv, seed_r, seed_scale, cand_key, cand_pos, revive, first_corr,
                       first_key, C, cs, dup, 0, split->after, split->open_mask, (const uint32_t*)nullptr,
                       (const uint64_t*)nullptr, (int64_t)0);
    hipLaunchKernelGGL(k_open_count, dim3(nblocks(P + 1)), dim3(kBlock), 0, s, split->open_mask, P, split->cnt);
    exclusive_scan_u32(split->cnt, split->off, P + 1, split->tmp, s);
    hipLaunchKernelGGL(k_open_fill, dim3((unsigned)P), dim3(64), 0, s, split->open_mask, split->off, split->list);
    hipLaunchKernelGGL(kern, dim3(g), dim3(64), 0, s, pts, nodes, cl, ncl, bucket, lt, n,
                       (const int32_t*)split->list, nq, label, check_labels ? 1 : 0, inv, seed_r, seed_scale,
                       cand_key, cand_pos, revive, first_corr, first_key, C, cs, dup, split->after + 1, kMaxCappedPasses,
                       (uint64_t*)nullptr, (const uint32_t*)(split->off + P), (const uint64_t*)split->open_mask,
                       (int64_t)(split->compact_frac * (double)nq));
}

void launch_leaves(const uint64_t* keys, int64_t n, Cloud* cl, int ncl, int bucket, uint32_t* mark,
                   uint32_t* idx, uint32_t* tmp, int32_t* start, int32_t* of, int32_t* cloud_lf, const Params* P,
                   Counters* C, bool layout, hipStream_t s)
{
    if (n > 0) {
        hipLaunchKernelGGL(k_leaf_marks, dim3(nblocks(n)), dim3(kBlock), 0, s, keys, n, (const Cloud*)cl, ncl,
                           bucket, mark);
        exclusive_scan_u32(mark, idx, n, tmp, s);
        hipLaunchKernelGGL(k_leaf_fill, dim3(nblocks(n)), dim3(kBlock), 0, s, (const uint32_t*)mark,
                           (const uint32_t*)idx, n, start, of);
    }
    hipLaunchKernelGGL(k_cloud_leaves, dim3(1), dim3(1024), 0, s, cl, ncl, (const uint32_t*)mark,
                       (const uint32_t*)idx, n, cloud_lf, P, C, layout ? 1 : 0);
}

void launch_zero_u32(uint32_t* p, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n);
}

void launch_node_min_ids(const float4* pts, const Cloud* cl, int ncl, int64_t nleaves, LeafTable lt, Node* nodes,
                         unsigned long long* visit, hipStream_t s)
{
    if (nleaves <= 0) return;
    hipLaunchKernelGGL(k_nmin_tree, dim3(nblocks(nleaves)), dim3(kBlock), 0, s, pts, cl, ncl, nleaves, lt, nodes,
                       visit);
}

void launch_comp_seed_init(int64_t n, const float4* pts, int32_t* csize, uint32_t* cbox, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_comp_seed_init, dim3(nblocks(n)), dim3(kBlock), 0, s, n, pts, csize, cbox);
}

void launch_comp_round(int dim, int64_t n, const float4* pts, const uint32_t* ids, Node* nodes,
                       unsigned long long* visit,
                       const Cloud* cl, int ncl, int64_t nleaves,
                       int bucket, LeafTable lt,
                       const int32_t* label, int32_t* corr, uint64_t* corr_key, uint64_t* comp1, const int32_t* inv,
                       float seed_scale, const SplitScratch& sc, const CompSeed& cs, const NbrList& nb,
                       const int32_t* eu, const int32_t* eround, Counters* C, hipStream_t s, hipEvent_t* ev)
{
    if (n <= 0) return;
    const unsigned g = nblocks(n);
    if (cs.features_on)
        hipLaunchKernelGGL(k_comp_merge, dim3(g), dim3(kBlock), 0, s, n, label, eu, eround, cs.csize, cs.cbox,
                           (const Counters*)C);
    hipLaunchKernelGGL(k_nlab_tree, dim3(nblocks(nleaves)), dim3(kBlock), 0, s, label, cl, ncl, nleaves, lt,
                       nodes, visit, (const Counters*)C);
    const int64_t P = (n + 63) / 64;
    hipLaunchKernelGGL(dim == 3 ? k_comp_keep<kKnn3> : k_comp_keep<kKnn2>, dim3(g), dim3(kBlock), 0, s, n, ids,
                       label, corr, corr_key, comp1, nb,
                       (const Counters*)C);
    hipLaunchKernelGGL(k_comp_prep, dim3(g), dim3(kBlock), 0, s, n, pts, label, corr, (const uint64_t*)corr_key,
                       (const uint64_t*)comp1, sc.open_mask, (const Counters*)C);
    open_compact(sc.open_mask, P, sc.off, sc.tmp, nullptr, sc.list, s);
    const uint32_t* nq = sc.off + P;
    const float* seed_r = nullptr;
    if (cs.features_on) {
        hipLaunchKernelGGL(k_comp_features, dim3(g), dim3(kBlock), 0, s, cs.samples != nullptr ? 1 : 0,
                           cs.model != nullptr ? 1 : 0, dim, pts, nodes, cl, ncl, lt,
                           (const int32_t*)sc.list, nq, label, (const uint64_t*)corr_key, (const uint64_t*)comp1,
                           (const int32_t*)cs.csize, (const uint32_t*)cs.cbox, (const Counters*)C, cs.X, cs.LS,
                           cs.nq_keep);
        if (cs.model != nullptr) {
            launch_mlp_infer_dev(cs.model, cs.X, cs.nq_keep + 1, n, cs.Y, s);
            hipLaunchKernelGGL(k_comp_seed, dim3(g), dim3(kBlock), 0, s, (const int32_t*)sc.list,
                               (const uint32_t*)(cs.nq_keep + 1), (const float*)cs.LS, (const float*)cs.Y, cs.R,
                               (const Counters*)C);
            seed_r = cs.R;
        }
    }
    auto kern = dim == 3 ? (bucket == 64 ? k_comp_search<3, 64> : k_comp_search<3, 0>)
                         : (bucket == 16 ? k_comp_search<2, 16> : k_comp_search<2, 0>);
    // (2D at its default 64-point leaves runs the runtime-size instance: a
    // <2, 64> instance measured 7% slower on the 2D power-law cloud)
    // pass 0 over the searching points; their open packets' lanes into open_mask
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(kern, dim3((unsigned)std::min<int64_t>(P, kSearchBlocks)), dim3(64), 0, s, pts,
                       (const Node*)nodes, cl, ncl, bucket, lt,
                       (const int32_t*)sc.list, nq, label, inv, seed_scale, corr, corr_key, comp1, 0, 0,
                       sc.open_mask, seed_r, sc.defer, C, ev != nullptr ? 1 : 0);
    if (ev) (void)hipEventRecord(ev[1], s);
    open_compact(sc.open_mask, P, sc.off, sc.tmp, (const int32_t*)sc.list, sc.list2, s);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(kern, dim3((unsigned)std::min<int64_t>(P, kResumeBlocks)), dim3(64), 0, s, pts,
                       (const Node*)nodes, cl, ncl, bucket, lt,
                       (const int32_t*)sc.list2, (const uint32_t*)(sc.off + P), label, inv, seed_scale, corr,
                       corr_key, comp1, 1, kMaxCappedPasses, (uint64_t*)nullptr, seed_r, sc.defer, C,
                       ev != nullptr ? 1 : 0);
    // the queries the packets handed over, one wave each
    if (sc.defer != nullptr) {
        auto dk = dim == 3 ? k_comp_search_deferred<3> : k_comp_search_deferred<2>;
        hipLaunchKernelGGL(dk, dim3(kDeferBlocks), dim3(64 * kDeferWaves), 0, s, pts, (const Node*)nodes, cl, ncl, lt,
                           (const uint64_t*)sc.defer, label, inv, seed_scale, corr, corr_key, comp1, seed_r, C,
                           ev != nullptr ? 1 : 0);
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    // comp1 now holds every component's edge: the sample log of this round
    if (cs.features_on && cs.samples != nullptr)
        hipLaunchKernelGGL(k_comp_samples, dim3(g), dim3(kBlock), 0, s, (const int32_t*)sc.list,
                           (const uint32_t*)cs.nq_keep, (const float*)cs.X, (const float*)cs.LS, label,
                           (const int32_t*)corr, (const uint64_t*)corr_key, (const uint64_t*)comp1, cs.per_round,
                           cs.cap, cs.samples, C);
}

void launch_knn_first(int dim, int64_t n, const float4* pts, const Node* nodes, const Cloud* cl, int ncl,
                      int bucket, LeafTable lt, const int32_t* inv, int32_t* corr, uint64_t* corr_key,
                      const NbrList& nb, Counters* C, bool collect, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned P = (unsigned)((n + 63) / 64);
    const int cs = collect ? 1 : 0;
    if (dim == 3)
        hipLaunchKernelGGL(k_knn_first<3>, dim3(P), dim3(64), 0, s, n, pts, nodes, cl, ncl, bucket, lt, inv, corr,
                           corr_key, nb, C, cs);
    else
        hipLaunchKernelGGL(k_knn_first<2>, dim3(P), dim3(64), 0, s, n, pts, nodes, cl, ncl, bucket, lt, inv, corr,
                           corr_key, nb, C, cs);
}

void launch_dup_fill(int dim, int64_t n, const int32_t* label, const float4* pts, const int32_t* inv,
                     uint64_t* cand_key, int32_t* cand_pos, int32_t* revive, int32_t* first_corr,
                     uint64_t* first_key, Counters* C, DupState dup, hipStream_t s)
{
    if (n <= 0 || dup.rep == nullptr || dup.npoints <= 0) return;
    hipLaunchKernelGGL(k_dup_reset, dim3((unsigned)std::min<int64_t>(nblocks(dup.ngroups), 256)), dim3(kBlock), 0,
                       s, C, dup);
    const unsigned g = (unsigned)std::min<int64_t>(nblocks(dup.npoints), 1024);
    hipLaunchKernelGGL(k_dup_m2, dim3(g), dim3(kBlock), 0, s, label, pts, C, dup);
    if (dim == 3)
        hipLaunchKernelGGL(k_dup_fill<3>, dim3(g), dim3(kBlock), 0, s, n, label, pts, inv, cand_key, cand_pos,
                           revive, first_corr, first_key, C, dup);
    else
        hipLaunchKernelGGL(k_dup_fill<2>, dim3(g), dim3(kBlock), 0, s, n, label, pts, inv, cand_key, cand_pos,
                           revive, first_corr, first_key, C, dup);
}

void launch_round_init(int64_t n, int32_t* label, int32_t* label2, int32_t* eu, int32_t* corr,
                       uint64_t* corr_key, uint64_t* comp1, uint32_t* comp2, Counters* C, uint32_t target,
                       uint32_t max_rounds, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_init, dim3(nblocks(std::max<int64_t>(n, kWorkKinds * kWorkSlots))), dim3(kBlock), 0, s, n, label,
                       label2, eu, corr, corr_key, comp1, comp2, C, target, max_rounds);
}

void launch_round_begin(Counters* C, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(1), 0, s, C);
}

void launch_round_end(Counters* C, bool grouped, hipStream_t s, unsigned long long* stamps, uint32_t cap)
{
    hipLaunchKernelGGL(k_round_end, dim3(1), dim3(1), 0, s, C, grouped ? 1 : 0, stamps, cap);
}

void launch_stamp(unsigned long long* at, hipStream_t s)
{
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, at);
}

void launch_flag_research(int K, int64_t n, const uint64_t* cand_key, int32_t* list, Counters* C,
                          hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_flag_research, dim3(nblocks(n)), dim3(kBlock), 0, s, K, n, cand_key, list, C);
}

void launch_rank_cones(int dim, int64_t n, const uint64_t* cand_key, const int32_t* cand_pos,
                       const int32_t* qlist, int64_t nq, uint64_t* rkey, int32_t* rpos, uint8_t* cursor,
                       const Counters* C, hipStream_t s)
{
    if (nq <= 0) return;
    hipLaunchKernelGGL(dim == 3 ? k_rank_cones<3> : k_rank_cones<2>, dim3(nblocks(nq)), dim3(kBlock), 0, s, n,
                       cand_key, cand_pos, qlist, nq, rkey, rpos, cursor, C);
}

void launch_round_phase(int phase, int K, int64_t n, const uint32_t* ids, const uint64_t* rkey,
                        const int32_t* rpos, uint8_t* cursor, int32_t* label,
                        int32_t* label2, int32_t* corr,
                        uint64_t* corr_key, uint64_t* comp1, uint32_t* comp2,
                        Counters* C, int32_t* eu, int32_t* ev, float* ed2, int32_t* eround,
                        const CycleState& cy, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = nblocks(n);
    switch (phase) {
    case 0:
        hipLaunchKernelGGL(k_round_min, dim3(g), dim3(kBlock), 0, s, K, n, ids, rkey, rpos, cursor,
                           label, corr, corr_key, comp1, C);
        break;
    case 1:
        hipLaunchKernelGGL(k_round_comp2, dim3(g), dim3(kBlock), 0, s, n, ids, label, corr, corr_key, comp1,
                           comp2, C);
        break;
    case 2:
        if (cy.j0) {
            const unsigned gc = (unsigned)std::min<int64_t>(g, 1024);
            hipLaunchKernelGGL(k_cycle_init, dim3((unsigned)((n + kCyChunk - 1) / kCyChunk)), dim3(kBlock), 0, s, n,
                               label, comp1, comp2, cy, C);
            for (int k = 0; k < cy.steps; ++k)
                hipLaunchKernelGGL(k_cycle_step, dim3(gc), dim3(kBlock), 0, s, cy, k, C);
            hipLaunchKernelGGL(k_cycle_mark, dim3(gc), dim3(kBlock), 0, s, cy, C);
            hipLaunchKernelGGL(k_cycle_cut, dim3(gc), dim3(kBlock), 0, s, comp1, comp2, cy, C);
        }
        hipLaunchKernelGGL(k_round_hook, dim3(nblocks(n, kHookBlock)), dim3(kHookBlock), 0, s, n, ids, label,
                           label2, corr, corr_key, comp1, comp2, C, eu, ev, ed2, eround, cy.j0 ? cy.cut : nullptr);
        break;
    case 3:
        hipLaunchKernelGGL(k_round_flatten, dim3(g), dim3(kBlock), 0, s, n, label, label2, comp1, comp2, C);
        break;
    }
}

void launch_weight(const float* xyz, int dim, const int32_t* eu, const int32_t* ev, int64_t m,
                   double* partial, int nblk, hipStream_t s)
{
    if (dim == 3) hipLaunchKernelGGL(k_weight<3>, dim3(nblk), dim3(kBlock), 0, s, xyz, eu, ev, m, partial);
    else hipLaunchKernelGGL(k_weight<2>, dim3(nblk), dim3(kBlock), 0, s, xyz, eu, ev, m, partial);
}

void launch_cloud_results(const float* xyz, int dim, const Cloud* cl, int ncl, const int32_t* eu,
                          const int32_t* ev, const int32_t* eround, double* weight, int32_t* edges,
                          int32_t* last_round, hipStream_t s)
{
    if (ncl <= 0) return;
    if (dim == 3)
        hipLaunchKernelGGL(k_cloud_results<3>, dim3(ncl), dim3(kBlock), 0, s, xyz, cl, eu, ev, eround, weight,
                           edges, last_round);
    else
        hipLaunchKernelGGL(k_cloud_results<2>, dim3(ncl), dim3(kBlock), 0, s, xyz, cl, eu, ev, eround, weight,
                           edges, last_round);
}

void launch_labels_out(int64_t n, const int32_t* label, const float4* pts, const int32_t* inv,
                       int32_t* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_labels_out, dim3(nblocks(n)), dim3(kBlock), 0, s, n, label, pts, inv, out);
}

void launch_cands_out(int K, int64_t n, const uint64_t* cand_key, const int32_t* inv, int32_t* cand,
                      float* cd2, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cands_out, dim3(nblocks(n)), dim3(kBlock), 0, s, K, n, cand_key, inv, cand, cd2);
}

void launch_edge_keys(const int32_t* eu, const int32_t* ev, int64_t m, uint64_t* keys, hipStream_t s)
{
    if (m <= 0) return;
    hipLaunchKernelGGL(k_edge_keys, dim3(nblocks(m)), dim3(kBlock), 0, s, eu, ev, m, keys);
}

void launch_edges_out(const uint64_t* keys, const uint32_t* idx, const float* ed2, int64_t m, int32_t* u,
                      int32_t* v, float* d2, hipStream_t s)
{
    if (m <= 0) return;
    hipLaunchKernelGGL(k_edges_out, dim3(nblocks(m)), dim3(kBlock), 0, s, keys, idx, ed2, m, u, v, d2);
}

void launch_cell_grid(const float* resc, int64_t n, int dim, int pstride, const int32_t* geo, int32_t* cell_of,
                      uint64_t* keys, uint32_t* cell_count, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_grid, dim3(nblocks(n)), dim3(kBlock), 0, s, resc, n, dim, pstride, geo, cell_of,
                       keys, cell_count);
}

}  // namespace emst
