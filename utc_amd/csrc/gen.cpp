// gen.cpp -- seeded input generators for the benchmark and test systems.
//
// The reference builds its inputs with glibc rand() (benchapps/nbody/seq/nbody.h:89-250,
// restated by the UTC driver's NbodyInit task); those stay in the host
// driver.  The generators here produce the uniform-cube, log-uniform-mass,
// Plummer and target inputs, which have no reference counterpart, from one
// defined random stream so the library, the tests and the benchmark all see
// identical inputs:
//   * splitmix64 (Steele, Lea, Flood 2014) expands a 64-bit seed,
//   * xoshiro256** (Blackman, Vigna 2018) is the stream,
//   * uniforms are the top 53 bits times 2^-53,
//   * normals come from the Box-Muller transform, both outputs used,
//   * the Plummer sphere follows Aarseth, Henon & Wielen (1974): radius from
//     the cumulative mass M(r) = r^3/(1+r^2)^{3/2} with mass fraction cut at
//     0.999, isotropic direction, speed q*v_esc with q drawn by von Neumann
//     rejection from q^2 (1-q^2)^{7/2} (bound 0.1), v_esc = sqrt(2)(1+r^2)^{-1/4}
//     (G = M = a = 1); centre of mass and mean velocity removed.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "utcnb.h"

namespace utcnb {
int fail(int code, const std::string& msg);  // capi.cpp: records utcnb_last_error()
}

namespace {

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Stream {
    uint64_t s[4];
    bool have_spare = false;
    double spare = 0.0;
    explicit Stream(uint64_t seed) { utcnb_rng_seed(seed, s); }
    uint64_t next() { return utcnb_rng_next(s); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal()
    {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        double u1 = 1.0 - uniform();  // (0, 1]
        double u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        double t = 2.0 * M_PI * u2;
        spare = r * std::sin(t);
        have_spare = true;
        return r * std::cos(t);
    }
};

void isotropic(Stream& g, double len, double* out)
{
    double z = (1.0 - 2.0 * g.uniform()) * len;
    double phi = 2.0 * M_PI * g.uniform();
    double rxy = std::sqrt(std::fmax(len * len - z * z, 0.0));
    out[0] = rxy * std::cos(phi);
    out[1] = rxy * std::sin(phi);
    out[2] = z;
}

}  // namespace

extern "C" {

uint64_t utcnb_splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void utcnb_rng_seed(uint64_t seed, uint64_t state[4])
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i)
        state[i] = utcnb_splitmix64(&x);
}

uint64_t utcnb_rng_next(uint64_t s[4])
{
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

int utcnb_generate(int kind, int n, uint64_t seed, double p0, double p1, double p2,
                   double* pos4, double* vel4, double* out3)
{
    if (n <= 0)
        return utcnb::fail(UTCNB_EINVAL, "utcnb_generate: n must be >= 1");
    Stream g(seed);
    if (kind == UTCNB_GEN_TARGETS) {
        if (!out3)
            return utcnb::fail(UTCNB_EINVAL, "utcnb_generate: targets need out3");
        for (int64_t i = 0; i < 3 * (int64_t)n; ++i)
            out3[i] = p0 * g.normal();
        return UTCNB_OK;
    }
    if (!pos4 || !vel4)
        return utcnb::fail(UTCNB_EINVAL, "utcnb_generate: pos4 and vel4 are required");
    if (kind == UTCNB_GEN_CUBE || kind == UTCNB_GEN_CUBE_LOGM) {
        if (kind == UTCNB_GEN_CUBE_LOGM && !(p1 > 0 && p2 >= p1))
            return utcnb::fail(UTCNB_EINVAL, "utcnb_generate: log-uniform masses need 0 < p1 <= p2");
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < 3; ++k)
                pos4[4 * (size_t)i + k] = 2.0 * g.uniform() - 1.0;
            pos4[4 * (size_t)i + 3] = 1.0 / n;
        }
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < 3; ++k)
                vel4[4 * (size_t)i + k] = p0 * g.normal();
            vel4[4 * (size_t)i + 3] = 1.0;
        }
        if (kind == UTCNB_GEN_CUBE_LOGM) {
            const double l1 = std::log(p1), l2 = std::log(p2);
            for (int i = 0; i < n; ++i)
                pos4[4 * (size_t)i + 3] = std::exp(l1 + (l2 - l1) * g.uniform()) / n;
        }
        return UTCNB_OK;
    }
    if (kind == UTCNB_GEN_PLUMMER) {
        double cm[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
        for (int i = 0; i < n; ++i) {
            double x1;
            do {
                x1 = g.uniform();
            } while (x1 <= 0.0 || x1 > 0.999);
            double r = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
            double* p = pos4 + 4 * (size_t)i;
            isotropic(g, r, p);
            p[3] = 1.0 / n;
            double q, y;
            do {
                q = g.uniform();
                y = 0.1 * g.uniform();
            } while (y >= q * q * std::pow(1.0 - q * q, 3.5));
            double ve = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
            double* v = vel4 + 4 * (size_t)i;
            isotropic(g, q * ve, v);
            v[3] = 1.0;
            for (int k = 0; k < 3; ++k) {
                cm[k] += p[k];
                cv[k] += v[k];
            }
        }
        for (int k = 0; k < 3; ++k) {
            cm[k] /= n;
            cv[k] /= n;
        }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                pos4[4 * (size_t)i + k] -= cm[k];
                vel4[4 * (size_t)i + k] -= cv[k];
            }
        return UTCNB_OK;
    }
    return utcnb::fail(UTCNB_EINVAL, "utcnb_generate: unknown kind");
}

}  // extern "C"
