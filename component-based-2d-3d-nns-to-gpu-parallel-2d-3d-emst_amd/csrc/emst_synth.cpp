// emst_synth.cpp -- synthetic point clouds, identical on every host.
//
// SplitMix64 expands the seed into the 128-bit state of xoshiro128+; uniforms
// are u = (x >> 8) * 2^-24 in [0, 1).  Gaussians use Box-Muller in double.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/emst.h"

namespace {

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro128p {
    uint32_t s[4];
    explicit Xoshiro128p(uint64_t seed)
    {
        uint64_t sm = seed;
        const uint64_t a = splitmix64(sm), b = splitmix64(sm);
        s[0] = (uint32_t)a; s[1] = (uint32_t)(a >> 32);
        s[2] = (uint32_t)b; s[3] = (uint32_t)(b >> 32);
        if ((s[0] | s[1] | s[2] | s[3]) == 0) s[0] = 1;
    }
    static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    uint32_t next()
    {
        const uint32_t r = s[0] + s[3];
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return r;
    }
    float u() { return (float)(next() >> 8) * 0x1p-24f; }
    double gauss()
    {
        const double u1 = 1.0 - (double)u();     // (0, 1]
        const double u2 = (double)u();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

float clamp01(double v)
{
    const float hi = std::nextafter(1.0f, 0.0f);
    float f = (float)v;
    if (!(f >= 0.f)) f = 0.f;
    if (f > hi) f = hi;
    return f;
}

}  // namespace

extern "C" int emst_synth(int32_t kind, int64_t n, int32_t dim, uint64_t seed, float* out)
{
    if (n < 0 || (n > 0 && !out) || (dim != 2 && dim != 3)) return EMST_E_ARG;
    Xoshiro128p g(seed);
    if (kind == 0) {
        for (int64_t i = 0; i < n * dim; ++i) out[i] = g.u();
        return EMST_OK;
    }
    if (kind == 1 || kind == 3) {
        // kind 1: 64 isotropic Gaussians, sigma log-uniform in [0.002, 0.02],
        // 10% uniform background; kind 3: 8 such Gaussians, no background
        const int nc = kind == 1 ? 64 : 8;
        const float bg = kind == 1 ? 0.1f : 0.f;
        double cen[64][3], sig[64];
        for (int c = 0; c < nc; ++c) {
            for (int k = 0; k < dim; ++k) cen[c][k] = g.u();
            const double lo = std::log(0.002), hi = std::log(0.02);
            sig[c] = std::exp(lo + (double)g.u() * (hi - lo));
        }
        for (int64_t i = 0; i < n; ++i) {
            float* p = out + i * dim;
            if (g.u() < bg) {
                for (int k = 0; k < dim; ++k) p[k] = g.u();
            } else {
                int c = (int)(g.u() * nc);
                if (c >= nc) c = nc - 1;
                for (int k = 0; k < dim; ++k) p[k] = clamp01(cen[c][k] + sig[c] * g.gauss());
            }
        }
        return EMST_OK;
    }
    if (kind == 2) {
        // 16 centres, radial density ~ r^-1.5 on r in [1e-4, 0.5]
        const int nc = 16;
        double cen[16][3];
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < dim; ++k) cen[c][k] = g.u();
        const double rmin = 1e-4, rmax = 0.5;
        for (int64_t i = 0; i < n; ++i) {
            float* p = out + i * dim;
            int c = (int)(g.u() * nc);
            if (c >= nc) c = nc - 1;
            const double u = (double)g.u();
            if (dim == 2) {
                // radial pdf ~ r * r^-1.5 = r^-0.5: r = (sqrt(rmin) + u (sqrt(rmax) - sqrt(rmin)))^2
                const double sr = std::sqrt(rmin) + u * (std::sqrt(rmax) - std::sqrt(rmin));
                const double r = sr * sr;
                const double th = 2.0 * M_PI * (double)g.u();
                p[0] = clamp01(cen[c][0] + r * std::cos(th));
                p[1] = clamp01(cen[c][1] + r * std::sin(th));
            } else {
                // radial pdf ~ r^2 * r^-1.5 = r^0.5: r = (rmin^1.5 + u (rmax^1.5 - rmin^1.5))^(2/3)
                const double a = std::pow(rmin, 1.5), b = std::pow(rmax, 1.5);
                const double r = std::pow(a + u * (b - a), 2.0 / 3.0);
                const double z = 2.0 * (double)g.u() - 1.0;
                const double ph = 2.0 * M_PI * (double)g.u();
                const double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
                p[0] = clamp01(cen[c][0] + r * rr * std::cos(ph));
                p[1] = clamp01(cen[c][1] + r * rr * std::sin(ph));
                p[2] = clamp01(cen[c][2] + r * z);
            }
        }
        return EMST_OK;
    }
    return EMST_E_ARG;
}
