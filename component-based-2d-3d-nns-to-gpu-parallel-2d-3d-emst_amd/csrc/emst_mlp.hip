// emst_mlp.hip -- learned seed radius for the cone search.
//
// No counterpart in the reference (whose ring search always starts at its
// remembered d_min, NodeEMST.h:209-214): a small MLP 8 -> 64 -> 64 -> 1 predicts
// the log of the radius within which all of a point's cones close, relative to
// its leaf cell's mean spacing.  The prediction only seeds the iterative
// deepening of k_cone_search, so the search result never depends on it.
//
// Arithmetic: BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_16x16x32_bf16 for the 64-deep products, 16x16x16 for the
// 16 padded features), FP32 master weights and Adam.
//
// Model layout (EMST_MLP_PARAMS floats): W1[8][64] | b1[64] | W2[64][64] |
// b2[64] | w3[64] | b3 | feature mean[8] | feature 1/std[8] | seed offset.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emst.h"
#include "emst_internal.h"

namespace emst {

namespace {

constexpr int kF = 8;        // features
constexpr int kFP = 16;      // features padded to one MFMA K step
constexpr int kH = 64;       // hidden width
constexpr int oW1 = 0, oB1 = oW1 + kF * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH, oMean = oB3 + 1, oIstd = oMean + kF, oOff = oIstd + kF,
              kParams = oOff + 1;
constexpr int kTrain = oMean;    // trainable parameters come first
constexpr float kXClamp = 4.f;   // normalised features are clamped to [-4, 4]
static_assert(kParams == EMST_MLP_PARAMS, "model layout");

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);      // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// One wave: C[MT*16][NT*16] += A[M][K] * B[K][N] from LDS bf16 with strides,
// K = 16 KT.  An even KT runs gfx950's 16x16x32 BF16 product (half the
// matrix instructions): lane l holds A(m = 16mt + l%16, k = 32kt + 8(l/16) + j)
// and B(k, n = 16nt + l%16), j = 0..7; KT = 1 (the 16 padded features) runs
// the 16x16x16 product, lane l holding k = 16kt + 4(l/16) + j, j = 0..3.  Both
// leave C(m = 16mt + 4(l/16) + j, n = 16nt + l%16), j = 0..3, in lane l.
template <int MT, int NT, int KT>
__device__ __forceinline__ void wave_gemm(const uint16_t* A, int sam, int sak, const uint16_t* B, int sbk,
                                          int sbn, f4 (&acc)[MT][NT])
{
    const int l = __lane_id();
    const int r = l & 15, q = l >> 4;
    if constexpr (KT % 2 == 0) {
#pragma unroll
        for (int kt = 0; kt < KT / 2; ++kt) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                bf8 a;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    a[j] = __builtin_bit_cast(__bf16, A[(mt * 16 + r) * sam + (kt * 32 + 8 * q + j) * sak]);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    bf8 b;
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        b[j] = __builtin_bit_cast(__bf16, B[(kt * 32 + 8 * q + j) * sbk + (nt * 16 + r) * sbn]);
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[mt][nt], 0, 0, 0);
                }
            }
        }
    } else {
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                s4 a;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    a[j] = (short)A[(mt * 16 + r) * sam + (kt * 16 + 4 * q + j) * sak];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    s4 b;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        b[j] = (short)B[(kt * 16 + 4 * q + j) * sbk + (nt * 16 + r) * sbn];
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt][nt], 0, 0, 0);
                }
            }
        }
    }
}

template <int MT, int NT>
__device__ __forceinline__ void zero(f4 (&acc)[MT][NT])
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
}

// Forward (and, when T != nullptr, backward) for 64 samples per block (one wave).
// X: m x 8 raw features (normalised here with the model's mean / 1/std and
// clamped to [-kXClamp, kXClamp], so a cloud unlike the training clouds
// cannot drive the network outside the range it was fitted on).
// Inference writes y (+ seed offset when add_offset).  Training accumulates
// the gradient of sum((y - t)^2) * gscale into G and the squared error into loss.
__global__ __launch_bounds__(64) void k_mlp(const float* __restrict__ X, const float* __restrict__ T,
                                            int64_t m, const float* __restrict__ P, float* __restrict__ Y,
                                            int add_offset, float* __restrict__ G, double* __restrict__ loss,
                                            float gscale)
{
    // 44 KB of LDS; transposed operands are read through strides
    __shared__ uint16_t sW1t[kH * kFP];     // [n][f]   W1 transposed, f padded
    __shared__ uint16_t sW2[kH * kH];       // [k][n]
    __shared__ uint16_t sXt[kFP * 64];      // [f][s]   normalised features
    __shared__ uint16_t sH1[64 * kH];       // [s][u]
    __shared__ uint16_t sH2[64 * kH];       // [s][u]
    __shared__ uint16_t sD2[64 * kH];       // [s][u]  dZ2
    __shared__ uint16_t sD1t[kH * 64];      // [u][s]  dZ1
    const int l = __lane_id();
    const int r = l & 15, q = l >> 4;
    const int64_t s0 = (int64_t)blockIdx.x * 64;
    const bool valid = s0 + l < m;

    for (int i = l; i < kH * kFP; i += 64) {
        const int n = i / kFP, f = i % kFP;
        sW1t[i] = f < kF ? f2bf(P[oW1 + f * kH + n]) : 0;
    }
    for (int i = l; i < kH * kH; i += 64) sW2[i] = f2bf(P[oW2 + i]);
    for (int f = 0; f < kFP; ++f) {
        float x = 0.f;
        if (f < kF && valid)
            x = fminf(fmaxf((X[(s0 + l) * kF + f] - P[oMean + f]) * P[oIstd + f], -kXClamp), kXClamp);
        sXt[f * 64 + l] = f2bf(x);
    }
    __syncthreads();

    // layer 1: Z1 = Xn W1 + b1; H1 = relu(Z1)
    f4 acc[4][4];
    zero(acc);
    wave_gemm<4, 4, 1>(sXt, 1, 64, sW1t, 1, kFP, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = mt * 16 + 4 * q + j, u = nt * 16 + r;
                sH1[s * kH + u] = f2bf(fmaxf(acc[mt][nt][j] + P[oB1 + u], 0.f));
            }
    __syncthreads();
    // layer 2: Z2 = H1 W2 + b2; H2 = relu(Z2)
    zero(acc);
    wave_gemm<4, 4, 4>(sH1, kH, 1, sW2, kH, 1, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = mt * 16 + 4 * q + j, u = nt * 16 + r;
                sH2[s * kH + u] = f2bf(fmaxf(acc[mt][nt][j] + P[oB2 + u], 0.f));
            }
    __syncthreads();
    // layer 3: y = H2 . w3 + b3 (bf16 operands, fp32 sum)
    float y = 0.f;
    for (int u = 0; u < kH; ++u) y = __builtin_fmaf(bf2f(sH2[l * kH + u]), bf2f(f2bf(P[oW3 + u])), y);
    y += P[oB3];
    if (!T) {
        if (valid) Y[s0 + l] = y + (add_offset ? P[oOff] : 0.f);
        return;
    }

    // ---- backward of sum (y - t)^2, scaled by gscale
    const float e = valid ? (y - T[s0 + l]) : 0.f;
    const float dy = valid ? 2.f * e * gscale : 0.f;
    {
        double se = (double)e * e;
        for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
        if (l == 0) atomicAdd(loss, se);
        float db3 = dy;
        for (int off = 32; off > 0; off >>= 1) db3 += __shfl_xor(db3, off);
        if (l == 0) atomicAdd(&G[oB3], db3);
    }
    __syncthreads();
    // dw3[u] = sum_s dy[s] H2[s][u]; dZ2[s][u] = dy[s] w3[u] [H2 > 0]
    for (int u = 0; u < kH; ++u) {
        const float h = bf2f(sH2[l * kH + u]);
        float g = dy * h;
        for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off);
        if (l == 0) atomicAdd(&G[oW3 + u], g);
        const float d = h > 0.f ? dy * bf2f(f2bf(P[oW3 + u])) : 0.f;
        sD2[l * kH + u] = f2bf(d);
    }
    __syncthreads();
    // db2[u] = sum_s dZ2[s][u]  (lane = u)
    {
        float g = 0.f;
        for (int s = 0; s < 64; ++s) g += bf2f(sD2[s * kH + l]);
        atomicAdd(&G[oB2 + l], g);
    }
    // dW2 = H1^T dZ2: M = k (input unit), K = samples, N = n (output unit)
    zero(acc);
    wave_gemm<4, 4, 4>(sH1, 1, kH, sD2, kH, 1, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = mt * 16 + 4 * q + j, n = nt * 16 + r;
                atomicAdd(&G[oW2 + k * kH + n], acc[mt][nt][j]);
            }
    // dH1 = dZ2 W2^T: M = samples, K = output unit, N = input unit; dZ1 = dH1 [H1 > 0]
    zero(acc);
    wave_gemm<4, 4, 4>(sD2, kH, 1, sW2, 1, kH, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = mt * 16 + 4 * q + j, u = nt * 16 + r;
                const float d = bf2f(sH1[s * kH + u]) > 0.f ? acc[mt][nt][j] : 0.f;
                sD1t[u * 64 + s] = f2bf(d);
            }
    __syncthreads();
    // db1[u] = sum_s dZ1[s][u]  (lane = u)
    {
        float g = 0.f;
        for (int s = 0; s < 64; ++s) g += bf2f(sD1t[l * 64 + s]);
        atomicAdd(&G[oB1 + l], g);
    }
    // dW1 = Xn^T dZ1: M = features (16, 8 real), K = samples, N = hidden
    f4 acc1[1][4];
    zero(acc1);
    wave_gemm<1, 4, 4>(sXt, 64, 1, sD1t, 1, 64, acc1);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * q + j, u = nt * 16 + r;
            if (f < kF) atomicAdd(&G[oW1 + f * kH + u], acc1[0][nt][j]);
        }
}

// Inference: 4 waves per block share the staged weights and loop over
// 64-sample tiles (same arithmetic as k_mlp's forward).
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ X, int64_t m,
                                                  const float* __restrict__ P, float* __restrict__ Y,
                                                  int add_offset, const uint32_t* __restrict__ m_dev)
{
    if (m_dev != nullptr) m = *m_dev;
    __shared__ uint16_t sW1t[kH * kFP];
    __shared__ uint16_t sW2[kH * kH];
    __shared__ uint16_t sXt[4][kFP * 64];
    __shared__ uint16_t sH[4][64 * kH];
    __shared__ float sb1[kH], sb2[kH], sw3[kH];
    const int tid = threadIdx.x, w = tid / 64;
    const int l = __lane_id();
    const int r = l & 15, q = l >> 4;
    for (int i = tid; i < kH * kFP; i += 256) {
        const int n = i / kFP, f = i % kFP;
        sW1t[i] = f < kF ? f2bf(P[oW1 + f * kH + n]) : 0;
    }
    for (int i = tid; i < kH * kH; i += 256) sW2[i] = f2bf(P[oW2 + i]);
    if (tid < kH) {
        sb1[tid] = P[oB1 + tid];
        sb2[tid] = P[oB2 + tid];
        sw3[tid] = bf2f(f2bf(P[oW3 + tid]));
    }
    float mean[kF], istd[kF];
    for (int f = 0; f < kF; ++f) { mean[f] = P[oMean + f]; istd[f] = P[oIstd + f]; }
    const float b3 = P[oB3] + (add_offset ? P[oOff] : 0.f);
    __syncthreads();
    uint16_t* xt = sXt[w];
    uint16_t* hh = sH[w];
    const int64_t ntiles = (m + 63) / 64;
    for (int64_t t = (int64_t)blockIdx.x * 4 + w; t < ntiles; t += (int64_t)gridDim.x * 4) {
        const int64_t s0 = t * 64;
        const bool valid = s0 + l < m;
        for (int f = 0; f < kFP; ++f) {
            float x = 0.f;
            if (f < kF && valid) x = fminf(fmaxf((X[(s0 + l) * kF + f] - mean[f]) * istd[f], -kXClamp), kXClamp);
            xt[f * 64 + l] = f2bf(x);
        }
        __builtin_amdgcn_wave_barrier();
        f4 acc[4][4];
        zero(acc);
        wave_gemm<4, 4, 1>(xt, 1, 64, sW1t, 1, kFP, acc);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int s = mt * 16 + 4 * q + j, u = nt * 16 + r;
                    hh[s * kH + u] = f2bf(fmaxf(acc[mt][nt][j] + sb1[u], 0.f));
                }
        __builtin_amdgcn_wave_barrier();
        zero(acc);
        wave_gemm<4, 4, 4>(hh, kH, 1, sW2, kH, 1, acc);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int s = mt * 16 + 4 * q + j, u = nt * 16 + r;
                    hh[s * kH + u] = f2bf(fmaxf(acc[mt][nt][j] + sb2[u], 0.f));
                }
        __builtin_amdgcn_wave_barrier();
        float y = 0.f;
        for (int u = 0; u < kH; ++u) y = __builtin_fmaf(bf2f(hh[l * kH + u]), sw3[u], y);
        if (valid) Y[s0 + l] = y + b3;
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_adam(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M1,
                       float* __restrict__ M2, int n, float lr, float b1, float b2, float eps, int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    G[i] = 0.f;
    const float m1 = b1 * M1[i] + (1.f - b1) * g;
    const float m2 = b2 * M2[i] + (1.f - b2) * g * g;
    M1[i] = m1;
    M2[i] = m2;
    const float mh = m1 / (1.f - powf(b1, (float)t));
    const float vh = m2 / (1.f - powf(b2, (float)t));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

// Features of every sorted point for the seed model, all free of the cloud's
// scale (ratios of lengths, fractions), and LS = log s0, the leaf spacing the
// target and the seed are measured in.  Neighbours are the 8 nearest sorted
// positions (Morton order).
__global__ __launch_bounds__(256) void k_radius_features(int dim, const float4* __restrict__ pts,
                                                        const Node* __restrict__ nodes, int64_t L,
                                                        LeafTable lt, int64_t n, float* __restrict__ X,
                                                        float* __restrict__ LS)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 me = pts[p];
    const float pc[3] = {me.x, me.y, me.z};
    float sp[4];
    const int64_t nl = lt.of[n - 1] + 1;        // one cloud: leaves 0 .. nl - 1
    int64_t node = L + lt.of[p];
    Node leaf{};
    for (int k = 0; k < 4; ++k) {
        const Node nd = nodes[node];
        if (k == 0) leaf = nd;
        // points under this node: its leaves' range (slots past the last leaf are empty)
        int64_t first = node, last = node;
        while (first < L) { first = 2 * first; last = 2 * last + 1; }
        const int64_t a = first - L < nl ? lt.start[first - L] : n;
        const int64_t b = last - L < nl ? lt.start[last - L + 1] : n;
        const float cnt = (float)max<int64_t>(b - a, 1);
        float vol = 1.f, ext = 0.f;
        for (int d = 0; d < dim; ++d) {
            const float e = fmaxf(nd.hi[d] - nd.lo[d], 0.f);
            vol *= e;
            ext = fmaxf(ext, e);
        }
        float s = vol > 0.f ? powf(vol / cnt, 1.f / dim) : ext / cnt;
        sp[k] = fmaxf(s, 1e-6f);
        if (node > 1) node >>= 1;
    }
    float emin = INFINITY, emax = 0.f, fd = INFINITY;
    for (int d = 0; d < dim; ++d) {
        const float e = fmaxf(leaf.hi[d] - leaf.lo[d], 0.f);
        emin = fminf(emin, e);
        emax = fmaxf(emax, e);
        fd = fminf(fd, fminf(pc[d] - leaf.lo[d], leaf.hi[d] - pc[d]));
    }
    // sorted neighbours: distances (smallest, median), one-sidedness
    float dd[8];
    float dsum = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
    int nd = 0;
#pragma unroll
    for (int o = -4; o <= 4; ++o) {
        if (o == 0) continue;
        const int64_t q = p + o;
        float dist = INFINITY;
        if (q >= 0 && q < n) {
            const float4 w = pts[q];
            const float dx = w.x - me.x, dy = w.y - me.y, dz = dim == 3 ? w.z - me.z : 0.f;
            dist = sqrtf(dx * dx + dy * dy + dz * dz);
            dsum += dist;
            vx += dx; vy += dy; vz += dz;
            ++nd;
        }
        dd[o < 0 ? o + 4 : o + 3] = dist;
    }
    // sort the 8 distances (missing neighbours at the array ends are +inf)
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 7 - i; ++j) {
            const float a = dd[j], b = dd[j + 1];
            dd[j] = fminf(a, b);
            dd[j + 1] = fmaxf(a, b);
        }
    const float dmin = dd[0];
    const float dmed = dd[(nd - 1) / 2];
    const float s0 = sp[0];
    float* x = X + p * 8;
    x[0] = logf(fmaxf(dmin, 1e-3f * s0) / s0);
    x[1] = logf(sp[1] / sp[0]);
    x[2] = logf(sp[2] / sp[1]);
    x[3] = logf(sp[3] / sp[2]);
    x[4] = logf(fmaxf(emax, 1e-6f) / fmaxf(emin, 1e-6f * fmaxf(emax, 1e-6f)));
    x[5] = fmaxf(fd, 0.f) / s0;
    x[6] = logf(fmaxf(dmed, 1e-3f * s0) / s0);
    x[7] = dsum > 0.f ? sqrtf(vx * vx + vy * vy + vz * vz) / dsum : 0.f;
    LS[p] = logf(s0);
}

// Target: log(R / s0), R = sqrt(largest non-empty cone best d2) of the point,
// s0 = exp(LS) its leaf spacing.  Points with no candidate get -20.
__global__ __launch_bounds__(256) void k_radius_targets(int K, int64_t n, const uint64_t* __restrict__ cand_key,
                                                       const float* __restrict__ LS, float* __restrict__ T)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float mx = -1.f;
    for (int k = 0; k < K; ++k) {
        const uint64_t key = cand_key[(int64_t)k * n + p];
        if (key != kEmptyKey) mx = fmaxf(mx, __uint_as_float((uint32_t)(key >> 32)));
    }
    const float r = mx > 0.f ? sqrtf(mx) : 0.f;
    T[p] = r > 0.f ? fmaxf(logf(r) - LS[p], -20.f) : -20.f;
}

// Seed radius per sorted point from the model: s0 * exp(y + offset).
__global__ __launch_bounds__(256) void k_radius_seed(int64_t n, const float* __restrict__ LS,
                                                    const float* __restrict__ Y, float* __restrict__ R)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    R[p] = expf(LS[p] + fminf(Y[p], 20.f));
}

void launch_radius_features(int dim, const float4* pts, const Node* nodes, int64_t L, LeafTable lt, int64_t n,
                            float* X, float* LS, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_radius_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dim, pts, nodes, L,
                       lt, n, X, LS);
}

void launch_radius_targets(int K, int64_t n, const uint64_t* cand_key, const float* LS, float* T, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_radius_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, K, n, cand_key, LS, T);
}

void launch_radius_seed(int64_t n, const float* model, const float* X, const float* LS, float* Y, float* R,
                        hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_mlp_infer, dim3(g), dim3(256), 0, s, X, n, model, Y, 1, (const uint32_t*)nullptr);
    hipLaunchKernelGGL(k_radius_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, LS, Y, R);
}

void launch_mlp_infer_dev(const float* model, const float* X, const uint32_t* m_dev, int64_t m_max, float* Y,
                          hipStream_t s)
{
    if (m_max <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((m_max + 255) / 256, 1024);
    hipLaunchKernelGGL(k_mlp_infer, dim3(g), dim3(256), 0, s, X, m_max, model, Y, 1, m_dev);
}

}  // namespace emst

using namespace emst;

namespace {

struct Dev {
    void* p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) { return hipMalloc(&p, b ? b : 1); }
};

#define MCHK(expr)                                  \
    do {                                            \
        if ((expr) != hipSuccess) {                 \
            (void)hipGetLastError();                \
            return EMST_E_HIP;                      \
        }                                           \
    } while (0)

uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int emst_mlp_eval(int32_t device, const float* model, const float* feat, int64_t m, float* out)
{
    if (!model || m < 0 || (m > 0 && (!feat || !out))) return EMST_E_ARG;
    if (m == 0) return EMST_OK;
    MCHK(hipSetDevice(device));
    Dev dP, dX, dY;
    MCHK(dP.alloc(sizeof(float) * kParams));
    MCHK(dX.alloc(sizeof(float) * m * kF));
    MCHK(dY.alloc(sizeof(float) * m));
    MCHK(hipMemcpy(dP.p, model, sizeof(float) * kParams, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(dX.p, feat, sizeof(float) * m * kF, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, 0, (const float*)dX.p,
                       (const float*)nullptr, m, (const float*)dP.p, (float*)dY.p, 0, (float*)nullptr,
                       (double*)nullptr, 0.f);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(out, dY.p, sizeof(float) * m, hipMemcpyDeviceToHost));
    return EMST_OK;
}

int emst_mlp_grad(int32_t device, const float* model, const float* feat, const float* target, int64_t m,
                  float* grad, double* loss)
{
    if (!model || !grad || !loss || m <= 0 || !feat || !target) return EMST_E_ARG;
    MCHK(hipSetDevice(device));
    Dev dP, dX, dT, dG, dL;
    MCHK(dP.alloc(sizeof(float) * kParams));
    MCHK(dX.alloc(sizeof(float) * m * kF));
    MCHK(dT.alloc(sizeof(float) * m));
    MCHK(dG.alloc(sizeof(float) * kParams));
    MCHK(dL.alloc(sizeof(double)));
    MCHK(hipMemcpy(dP.p, model, sizeof(float) * kParams, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(dX.p, feat, sizeof(float) * m * kF, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(dT.p, target, sizeof(float) * m, hipMemcpyHostToDevice));
    MCHK(hipMemset(dG.p, 0, sizeof(float) * kParams));
    MCHK(hipMemset(dL.p, 0, sizeof(double)));
    hipLaunchKernelGGL(k_mlp, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, 0, (const float*)dX.p,
                       (const float*)dT.p, m, (const float*)dP.p, (float*)nullptr, 0, (float*)dG.p,
                       (double*)dL.p, 1.f);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(grad, dG.p, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    MCHK(hipMemcpy(loss, dL.p, sizeof(double), hipMemcpyDeviceToHost));
    return EMST_OK;
}

int emst_mlp_train(int32_t device, const float* feat, const float* target, int64_t m, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float* model, float* loss_per_epoch)
{
    if (!feat || !target || !model || m < 64 || epochs <= 0 || batch < 64 || batch % 64) return EMST_E_ARG;
    MCHK(hipSetDevice(device));
    // initialisation: feature normalisation from the data, He-uniform weights
    std::vector<float> P(kParams, 0.f);
    for (int f = 0; f < kF; ++f) {
        double s = 0, s2 = 0;
        for (int64_t i = 0; i < m; ++i) { const double x = feat[i * kF + f]; s += x; s2 += x * x; }
        const double mean = s / m, var = std::max(s2 / m - mean * mean, 1e-12);
        P[oMean + f] = (float)mean;
        P[oIstd + f] = (float)(1.0 / std::sqrt(var));
    }
    uint64_t st = seed;
    auto uni = [&]() { st = mix64(st); return (float)((st >> 40) * 0x1p-24) * 2.f - 1.f; };
    for (int i = 0; i < kF * kH; ++i) P[oW1 + i] = uni() * std::sqrt(6.f / kF);
    for (int i = 0; i < kH * kH; ++i) P[oW2 + i] = uni() * std::sqrt(6.f / kH);
    for (int i = 0; i < kH; ++i) P[oW3 + i] = uni() * std::sqrt(3.f / kH);
    double tm = 0;
    for (int64_t i = 0; i < m; ++i) tm += target[i];
    P[oB3] = (float)(tm / m);
    P[oOff] = 0.f;
    Dev dP, dX, dT, dG, dM1, dM2, dL;
    const int64_t nb = m / batch;
    const int64_t mm = nb * batch;
    MCHK(dP.alloc(sizeof(float) * kParams));
    MCHK(dX.alloc(sizeof(float) * mm * kF));
    MCHK(dT.alloc(sizeof(float) * mm));
    MCHK(dG.alloc(sizeof(float) * kParams));
    MCHK(dM1.alloc(sizeof(float) * kParams));
    MCHK(dM2.alloc(sizeof(float) * kParams));
    MCHK(dL.alloc(sizeof(double)));
    MCHK(hipMemcpy(dP.p, P.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
    MCHK(hipMemset(dG.p, 0, sizeof(float) * kParams));
    MCHK(hipMemset(dM1.p, 0, sizeof(float) * kParams));
    MCHK(hipMemset(dM2.p, 0, sizeof(float) * kParams));
    std::vector<int64_t> perm(m);
    for (int64_t i = 0; i < m; ++i) perm[i] = i;
    std::vector<float> hx(mm * kF), ht(mm);
    int t = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = m - 1; i > 0; --i) {           // Fisher-Yates, seeded
            st = mix64(st);
            std::swap(perm[i], perm[(int64_t)(st % (uint64_t)(i + 1))]);
        }
        for (int64_t i = 0; i < mm; ++i) {
            std::memcpy(&hx[i * kF], &feat[perm[i] * kF], sizeof(float) * kF);
            ht[i] = target[perm[i]];
        }
        MCHK(hipMemcpy(dX.p, hx.data(), sizeof(float) * mm * kF, hipMemcpyHostToDevice));
        MCHK(hipMemcpy(dT.p, ht.data(), sizeof(float) * mm, hipMemcpyHostToDevice));
        MCHK(hipMemset(dL.p, 0, sizeof(double)));
        for (int64_t b = 0; b < nb; ++b) {
            hipLaunchKernelGGL(k_mlp, dim3((unsigned)(batch / 64)), dim3(64), 0, 0,
                               (const float*)dX.p + b * batch * kF, (const float*)dT.p + b * batch, (int64_t)batch,
                               (const float*)dP.p, (float*)nullptr, 0, (float*)dG.p, (double*)dL.p,
                               1.f / (float)batch);
            ++t;
            hipLaunchKernelGGL(k_adam, dim3((kTrain + 255) / 256), dim3(256), 0, 0, (float*)dP.p, (float*)dG.p,
                               (float*)dM1.p, (float*)dM2.p, kTrain, lr, 0.9f, 0.999f, 1e-8f, t);
        }
        MCHK(hipGetLastError());
        double L = 0;
        MCHK(hipMemcpy(&L, dL.p, sizeof(double), hipMemcpyDeviceToHost));
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(L / (double)mm);
    }
    MCHK(hipMemcpy(model, dP.p, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    return EMST_OK;
}

}  // extern "C"
