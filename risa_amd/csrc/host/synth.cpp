// Seeded synthetic frame source: random-ellipse phantoms, analytic fan and
// parallel line integrals, Poisson counts.  See include/risa_synth.h.
#include "risa_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// splitmix64 seeding + PCG-XSH-RR 32-bit generator (O'Neill, 2014).
uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t state, inc;
    explicit Rng(uint64_t seed, uint64_t stream = 0)
    {
        uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
        state = splitmix64(s);
        inc = splitmix64(s) | 1u;
        next();
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    // uniform in [0, 1) with 53 random bits
    double uniform()
    {
        const uint64_t hi = next(), lo = next();
        return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
    }
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    // Poisson variate: multiplication method below 10, PTRS (Hoermann 1993,
    // "The transformed rejection method for generating Poisson random
    // variables") above.
    long poisson(double lam)
    {
        if (lam <= 0.0) return 0;
        if (lam < 10.0) {
            const double L = std::exp(-lam);
            long k = 0;
            double p = uniform();
            while (p > L) { k++; p *= uniform(); }
            return k;
        }
        const double slam = std::sqrt(lam), loglam = std::log(lam);
        const double b = 0.931 + 2.53 * slam;
        const double a = -0.059 + 0.02483 * b;
        const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
        const double vr = 0.9277 - 3.6224 / (b - 2);
        for (;;) {
            const double U = uniform() - 0.5, V = uniform();
            const double us = 0.5 - std::fabs(U);
            const long k = (long)std::floor((2 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr) return k;
            if (k < 0 || (us < 0.013 && V > us)) continue;
            if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
                -lam + k * loglam - std::lgamma((double)k + 1))
                return k;
        }
    }
};

struct Ellipse { double cx, cy, a, b, phi, mu, c, s; };

std::vector<Ellipse> make_phantom(uint64_t seed, double width, const risa_synth_opts& o)
{
    Rng r(seed, 0x5EED);
    const int span = std::max(0, o.n_max - o.n_min);
    const int n = o.n_min + (int)(r.uniform() * (span + 1));
    std::vector<Ellipse> es;
    const double R = 0.5 * width;
    for (int k = 0; k < n; k++) {
        Ellipse e;
        const double rad = 0.8 * R * std::sqrt(r.uniform());
        const double ang = 2.0 * kPi * r.uniform();
        e.cx = rad * std::cos(ang);
        e.cy = rad * std::sin(ang);
        e.a = (0.05 + 0.35 * r.uniform()) * R;
        e.b = (0.05 + 0.35 * r.uniform()) * R;
        e.phi = kPi * r.uniform();
        e.mu = o.mu_max * r.uniform();
        e.c = std::cos(e.phi);
        e.s = std::sin(e.phi);
        es.push_back(e);
    }
    return es;
}

// Integral of the phantom along the line through p0 with unit direction u.
double line_integral(const std::vector<Ellipse>& es, double px, double py, double ux, double uy)
{
    double sum = 0.0;
    for (const Ellipse& e : es) {
        const double dx = px - e.cx, dy = py - e.cy;
        const double qx = (dx * e.c + dy * e.s) / e.a, qy = (-dx * e.s + dy * e.c) / e.b;
        const double vx = (ux * e.c + uy * e.s) / e.a, vy = (-ux * e.s + uy * e.c) / e.b;
        const double A = vx * vx + vy * vy, B = 2.0 * (qx * vx + qy * vy), C = qx * qx + qy * qy - 1.0;
        const double disc = B * B - 4.0 * A * C;
        if (disc > 0.0) sum += e.mu * std::sqrt(disc) / A;
    }
    return sum;
}

// The rebinning's source-ring map (cuda_kernels_fan2para.h:38-62) in double
// and its inverse by bisection: fan frames are generated with the in-plane
// source angle that the rebinning will map onto the fan angle theta_j.
double ring_map(double alpha, double DX, double DZ, double diam)
{
    const double L = std::sqrt(DX * DX + DZ * DZ), R = 0.5 * diam, CA = std::cos(alpha);
    const double clampe = std::min(1.0, std::max(-1.0, (L * L + R * DX * CA) / (L * std::sqrt(L * L + R * R + 2.0 * R * DX * CA))));
    const double eps = std::acos(clampe);
    const double p1 = (L * L - R * DX) / (L * std::sqrt(L * L + R * R - 2.0 * R * DX));
    const double p2 = (L * L + R * DX) / (L * std::sqrt(L * L + R * R + 2.0 * R * DX));
    const double gam = 0.5 * (std::acos(p1) - std::acos(p2));
    const double den = std::sqrt(eps * eps + 2.0 * eps * gam * CA + gam * gam);
    const double ae = den > 0 ? std::min(1.0, std::max(-1.0, (eps * CA + gam) / den)) : 1.0;
    return alpha <= kPi ? std::acos(ae) : 2.0 * kPi - std::acos(ae);
}

double ring_unmap(double target, double DX, double DZ, double diam)
{
    double lo = 0.0, hi = 2.0 * kPi;
    for (int it = 0; it < 60; it++) {
        const double mid = 0.5 * (lo + hi);
        if (ring_map(mid, DX, DZ, diam) < target) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

int thread_count(const risa_synth_opts& o)
{
    if (o.threads > 0) return o.threads;
    return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

template <class F>
void parallel_rows(int rows, int nth, F&& f)
{
    std::vector<std::thread> pool;
    const int chunk = (rows + nth - 1) / nth;
    for (int w = 0; w < nth; w++) {
        const int r0 = w * chunk, r1 = std::min(rows, r0 + chunk);
        if (r0 < r1) pool.emplace_back([&f, r0, r1] { for (int r = r0; r < r1; r++) f(r); });
    }
    for (auto& t : pool) t.join();
}

uint16_t clip16(double v)
{
    if (v < 0.0) return 0;
    if (v > 65535.0) return 65535;
    return (uint16_t)v;
}

}  // namespace

extern "C" {

void risa_synth_default_opts(risa_synth_opts* o)
{
    std::memset(o, 0, sizeof(*o));
    o->I0 = 60000.0;
    o->dark_mean = 200.0;
    o->dark_sd = 5.0;
    o->mu_max = 0.05;
    o->n_min = 10;
    o->n_max = 20;
}

int risa_synth_phantom(uint64_t seed, double width, const risa_synth_opts* o, double* out, int max_n)
{
    if (!o) return -1;
    const auto es = make_phantom(seed, width, *o);
    const int n = std::min<int>((int)es.size(), max_n);
    for (int k = 0; k < n; k++) {
        const Ellipse& e = es[k];
        const double v[6] = {e.cx, e.cy, e.a, e.b, e.phi, e.mu};
        std::memcpy(out + 6 * k, v, sizeof(v));
    }
    return n;
}

int risa_synth_phantom_image(uint64_t seed, double width, int N, const risa_synth_opts* o, float* img)
{
    if (!o || N <= 0) return -1;
    const auto es = make_phantom(seed, width, *o);
    const double h = width / N, c = (N - 1) * 0.5;
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            const double X = (x - c) * h, Y = (y - c) * h;
            double mu = 0.0;
            for (const Ellipse& e : es) {
                const double dx = X - e.cx, dy = Y - e.cy;
                const double u = (dx * e.c + dy * e.s) / e.a, v = (-dx * e.s + dy * e.c) / e.b;
                if (u * u + v * v <= 1.0) mu += e.mu;
            }
            img[x + (size_t)N * y] = (float)mu;
        }
    return 0;
}

int risa_synth_fan_integrals(const risa_geom* g, int plane, uint64_t seed, const risa_synth_opts* o, double* p)
{
    if (!g || !o || plane < 0 || plane >= g->numberOfPlanes) return -1;
    const int Df = g->numberOfFanDetectors, Pf = g->numberOfFanProjections;
    const auto es = make_phantom(seed, g->imageWidth, *o);
    const double rT = g->sourceDiameter[plane] / 2.0, rD = g->detectorDiameter / 2.0;
    std::vector<double> theta(Pf), gammaGeo(Df);
    for (int j = 0; j < Pf; j++) {
        double v = j * (360.0 / Pf) - g->sourceOffset;
        if (v < 0.0) v += 360.0;
        const double th = (2 * kPi / 360.0) * v;
        theta[j] = ring_unmap(th, g->deltaX[plane], g->deltaZ[plane], 2 * rT);
    }
    // detector i sits at angle gamma_i + pi/2 on the detector ring
    for (int i = 0; i < Df; i++) gammaGeo[i] = (2 * kPi / 360.0) * (i * (360.0 / Df)) + kPi / 2;
    parallel_rows(Pf, thread_count(*o), [&](int j) {
        const double sx = rT * std::cos(theta[j]), sy = rT * std::sin(theta[j]);
        for (int i = 0; i < Df; i++) {
            const double dx = rD * std::cos(gammaGeo[i]) - sx, dy = rD * std::sin(gammaGeo[i]) - sy;
            const double len = std::sqrt(dx * dx + dy * dy);
            p[(size_t)j * Df + i] = len > 0 ? line_integral(es, sx, sy, dx / len, dy / len) : 0.0;
        }
    });
    return 0;
}

int risa_synth_fan_frame(const risa_geom* g, int plane, uint64_t seed, uint64_t noise_seed,
                         const risa_synth_opts* o, uint16_t* raw)
{
    if (!g || !o) return -1;
    const int Df = g->numberOfFanDetectors, Pf = g->numberOfFanProjections;
    std::vector<double> p((size_t)Df * Pf);
    if (risa_synth_fan_integrals(g, plane, seed, o, p.data()) != 0) return -1;
    std::vector<uint8_t> dead(Df, 0);
    for (int k = 0; k < o->n_dead; k++)
        if (o->dead[k] >= 0 && o->dead[k] < Df) dead[o->dead[k]] = 1;
    parallel_rows(Pf, thread_count(*o), [&](int j) {
        Rng r(noise_seed, (uint64_t)j + 1);
        for (int i = 0; i < Df; i++) {
            const double darkv = std::round(o->dark_mean + o->dark_sd * r.normal());
            const double counts = dead[i] ? 0.0 : (double)r.poisson(o->I0 * std::exp(-p[(size_t)j * Df + i]));
            raw[(size_t)j * Df + i] = clip16(counts + darkv);
        }
    });
    return 0;
}

int risa_synth_reference_frame(const risa_geom* g, uint64_t noise_seed, const risa_synth_opts* o, uint16_t* raw)
{
    if (!g || !o) return -1;
    const int Df = g->numberOfFanDetectors, Pf = g->numberOfFanProjections;
    std::vector<uint8_t> dead(Df, 0);
    for (int k = 0; k < o->n_dead; k++)
        if (o->dead[k] >= 0 && o->dead[k] < Df) dead[o->dead[k]] = 1;
    parallel_rows(Pf, thread_count(*o), [&](int j) {
        Rng r(noise_seed, (uint64_t)j + 1);
        for (int i = 0; i < Df; i++) {
            const double darkv = std::round(o->dark_mean + o->dark_sd * r.normal());
            const double counts = dead[i] ? 0.0 : (double)r.poisson(o->I0);
            raw[(size_t)j * Df + i] = clip16(counts + darkv);
        }
    });
    return 0;
}

int risa_synth_parallel_sinogram(uint64_t seed, double width, int P, int D, double rotationOffset,
                                 uint64_t noise_seed, const risa_synth_opts* o, float* sino)
{
    if (!o || P <= 0 || D <= 0) return -1;
    const auto es = make_phantom(seed, width, *o);
    const double bin = width / D;
    const int c = D / 2;
    parallel_rows(P, thread_count(*o), [&](int pidx) {
        double th = pidx * kPi / P + rotationOffset / 180.0 * kPi;
        const double ct = std::cos(th), st = std::sin(th);
        Rng r(noise_seed, (uint64_t)pidx + 1);
        for (int k = 0; k < D; k++) {
            const double s = (k - c) * bin;
            double v = line_integral(es, s * ct, s * st, -st, ct);
            if (noise_seed != 0) {
                const double I = (double)clip16((double)r.poisson(o->I0 * std::exp(-v)));
                v = -std::log(std::max(I, 1.0) / o->I0);
            }
            sino[(size_t)pidx * D + k] = (float)v;
        }
    });
    return 0;
}

}  // extern "C"
