// Seeded synthetic initial conditions for the Barnes-Hut workloads (host code).
//
// The workload generators below follow the configurations the library is
// measured on: an exponential rotating disc, a uniform cold-collapse disc and a
// parabolic binary-disc encounter.  All are generated in f64 from a XorShift64
// stream (nbody/src/lib.rs:430-452) and rounded once to the f32 transfer
// records, which are then the canonical state.  gbh_ic_reference_disc /
// gbh_ic_reference_collision restate the reference presets make_disc /
// make_collision (nbody/src/lib.rs:454-528) in kpc / Msun / Myr units.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/galaxy_bh.h"

namespace gbh {
void set_error(const std::string& msg);
}

namespace {

constexpr double kTau = 6.283185307179586476925286766559;
constexpr double kGKpc = 4.498502151575286e-12;  // nbody/src/lib.rs:6
constexpr double kDefaultSofteningKpc = 0.05;    // nbody/src/lib.rs:8

// XorShift64 (nbody/src/lib.rs:430-452)
struct XorShift64 {
    uint64_t state;
    explicit XorShift64(uint64_t seed) : state(seed == 0 ? 1 : seed) {}
    uint64_t next() {
        uint64_t x = state;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        state = x;
        return x;
    }
    double unit() { return (double)(next() >> 11) / (double)(1ull << 53); }
    // Box-Muller, one deviate per call (cosine branch).
    double gauss() {
        double u1 = unit();
        double u2 = unit();
        double r = std::sqrt(-2.0 * std::log(u1 > 1e-300 ? u1 : 1e-300));
        return r * std::cos(kTau * u2);
    }
};

// Enclosed-mass fraction of an exponential disc at q = R / scale.
double exp_disc_cdf(double q) { return 1.0 - (1.0 + q) * std::exp(-q); }

// Starting points of the inverse below: q at targets k / kInvTable, k < kInvTable,
// each by bisection on [0, 64] (computed once per process).
constexpr int kInvTable = 4096;
const double* inverse_table() {
    static const std::vector<double> table = [] {
        std::vector<double> t(kInvTable + 1);
        for (int k = 0; k <= kInvTable; ++k) {
            const double target = (double)k / kInvTable;
            double lo = 0.0, hi = 64.0;
            for (int it = 0; it < 64; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (exp_disc_cdf(mid) < target) lo = mid; else hi = mid;
            }
            t[k] = 0.5 * (lo + hi);
        }
        return t;
    }();
    return table.data();
}

// Inverse of exp_disc_cdf on [0, qmax]: linear interpolation in the table, then
// Newton steps on cdf(q) - target (cdf'(q) = q e^-q), kept inside [0, qmax] by
// bisecting toward the violated end; a fixed sequence of operations per target,
// so the result depends only on the target and qmax.
double exp_disc_radius(double target, double qmax) {
    if (!(target > 0.0)) return 0.0;
    const double* tab = inverse_table();
    const double x = target * kInvTable;
    int k = (int)x;
    if (k >= kInvTable) k = kInvTable - 1;
    const double f = x - k;
    double q = tab[k] + f * (tab[k + 1] - tab[k]);
    if (!(q > 0.0)) q = std::sqrt(2.0 * target);  // cdf(q) ~ q^2 / 2 near 0
    if (q > qmax) q = qmax;
    for (int it = 0; it < 40; ++it) {
        const double e = std::exp(-q);
        const double r = 1.0 - (1.0 + q) * e - target;
        const double slope = q * e;
        if (!(slope > 0.0)) break;
        double next = q - r / slope;
        if (!(next > 0.0)) next = 0.5 * q;
        if (next > qmax) next = 0.5 * (q + qmax);
        const double step = next - q;
        q = next;
        if (std::fabs(step) <= 1e-15 * q) break;
    }
    return q;
}

bool store(GbhEvolveState* o, double x, double y, double m, double vx, double vy) {
    float fx = (float)x, fy = (float)y, fm = (float)m, fvx = (float)vx, fvy = (float)vy;
    if (!std::isfinite(fx) || !std::isfinite(fy) || !std::isfinite(fm) || !std::isfinite(fvx) ||
        !std::isfinite(fvy) || !(fm > 0.0f))
        return false;
    o->position_mass[0] = fx;
    o->position_mass[1] = fy;
    o->position_mass[2] = fm;
    o->position_mass[3] = 0.0f;
    o->velocity[0] = fvx;
    o->velocity[1] = fvy;
    o->velocity[2] = 0.0f;
    o->velocity[3] = 0.0f;
    return true;
}

int disc_into(uint32_t n, uint64_t seed, double total_mass, double scale, double truncation,
              double softening, double g, double dispersion, double cx, double cy, double cvx,
              double cvy, GbhEvolveState* out) {
    if (n < 1 || !(total_mass > 0) || !(scale > 0) || !(truncation > 0) || !(g > 0) ||
        !(softening >= 0) || !(dispersion >= 0)) {
        gbh::set_error("disc parameters must be positive and finite");
        return -1;
    }
    XorShift64 rng(seed);
    const double fmax = exp_disc_cdf(truncation);
    const double m = total_mass / (double)n;
    for (uint32_t i = 0; i < n; ++i) {
        double q = exp_disc_radius(rng.unit() * fmax, truncation);
        double r = scale * q;
        double phi = kTau * rng.unit();
        double c = std::cos(phi), s = std::sin(phi);
        double enclosed = exp_disc_cdf(q) / fmax;
        double soft_r = std::sqrt(r * r + softening * softening);
        if (soft_r < 1e-300) soft_r = 1e-300;
        double vc = std::sqrt(g * total_mass * enclosed / soft_r);
        double g1 = rng.gauss();
        double g2 = rng.gauss();
        double vx = -vc * s + dispersion * vc * g1;
        double vy = vc * c + dispersion * vc * g2;
        if (!store(out + i, cx + r * c, cy + r * s, m, cvx + vx, cvy + vy)) {
            gbh::set_error("disc body is not representable as finite f32");
            return -1;
        }
    }
    return 0;
}

// disc_body (nbody/src/lib.rs:454-483)
void reference_disc_body(XorShift64& rng, double per_body_mass, double total_mass, double scale,
                         double cx, double cy, double cvx, double cvy, double direction,
                         double spin, double softening, double* b) {
    double u = 1.0 - rng.unit();
    if (u < 1e-15) u = 1e-15;
    double r = scale * std::sqrt(-std::log(u));
    double angle = kTau * rng.unit();
    double x = cx + r * std::cos(angle);
    double y = cy + r * std::sin(angle);
    double q = r / scale;
    if (q < 0.0) q = 0.0;
    double enclosed = 1.0 - std::exp(-q) * (1.0 + q);
    if (enclosed < 0.0) enclosed = 0.0;
    if (enclosed > 1.0) enclosed = 1.0;
    double softened_r = std::sqrt(r * r + softening * softening);
    if (softened_r < 1e-9) softened_r = 1e-9;
    double circular = std::sqrt(kGKpc * total_mass * enclosed / softened_r);
    b[0] = x;
    b[1] = y;
    b[2] = cvx - direction * spin * circular * std::sin(angle);
    b[3] = cvy + direction * spin * circular * std::cos(angle);
    b[4] = per_body_mass;
}

}  // namespace

extern "C" {

int gbh_ic_rotating_disc(uint32_t n, uint64_t seed, double total_mass, double scale,
                         double truncation, double softening, double g, double dispersion,
                         GbhEvolveState* out) {
    return disc_into(n, seed, total_mass, scale, truncation, softening, g, dispersion, 0.0, 0.0,
                     0.0, 0.0, out);
}

int gbh_ic_uniform_collapse(uint32_t n, uint64_t seed, double total_mass, double radius,
                            GbhEvolveState* out) {
    if (n < 1 || !(total_mass > 0) || !(radius > 0)) {
        gbh::set_error("collapse parameters must be positive and finite");
        return -1;
    }
    XorShift64 rng(seed);
    const double m = total_mass / (double)n;
    for (uint32_t i = 0; i < n; ++i) {
        double r = radius * std::sqrt(rng.unit());
        double phi = kTau * rng.unit();
        if (!store(out + i, r * std::cos(phi), r * std::sin(phi), m, 0.0, 0.0)) {
            gbh::set_error("collapse body is not representable as finite f32");
            return -1;
        }
    }
    return 0;
}

int gbh_ic_binary_disc(uint32_t n1, uint32_t n2, uint64_t seed1, uint64_t seed2, double m1,
                       double m2, double separation, double pericentre, double softening,
                       double g, double dispersion, GbhEvolveState* out) {
    if (!(m1 > 0) || !(m2 > 0) || !(separation > 0) || !(pericentre > 0) ||
        pericentre > separation || !(g > 0)) {
        gbh::set_error("binary-disc parameters must be positive with pericentre <= separation");
        return -1;
    }
    const double mt = m1 + m2;
    const double v = std::sqrt(2.0 * g * mt / separation);         // parabolic speed at d
    const double l = std::sqrt(2.0 * g * mt * pericentre);         // specific angular momentum
    const double vt = l / separation;
    const double vr2 = v * v - vt * vt;
    const double vr = -std::sqrt(vr2 > 0 ? vr2 : 0.0);             // approaching
    // Disc 2 relative to disc 1: r = (separation, 0), v_rel = (vr, vt).
    const double f1 = m2 / mt, f2 = m1 / mt;
    if (disc_into(n1, seed1, m1, 1.0, 5.0, softening, g, dispersion, -f1 * separation, 0.0,
                  -f1 * vr, -f1 * vt, out) != 0)
        return -1;
    return disc_into(n2, seed2, m2, 1.0, 5.0, softening, g, dispersion, f2 * separation, 0.0,
                     f2 * vr, f2 * vt, out + n1);
}

// make_disc (nbody/src/lib.rs:485-502)
int gbh_ic_reference_disc(uint32_t n, uint64_t seed, double total_mass, double scale_kpc,
                          double* bodies) {
    if (n < 2 || n > 1000000) {
        gbh::set_error("particle count must be in 2..=1000000");
        return -1;
    }
    if (!std::isfinite(total_mass) || total_mass <= 0.0 || !std::isfinite(scale_kpc) ||
        scale_kpc <= 0.0) {
        gbh::set_error("disc mass and scale must be positive and finite");
        return -1;
    }
    XorShift64 rng(seed);
    double per = total_mass / (double)n;
    for (uint32_t i = 0; i < n; ++i)
        reference_disc_body(rng, per, total_mass, scale_kpc, 0.0, 0.0, 0.0, 0.0, 1.0, 0.88,
                            kDefaultSofteningKpc, bodies + 5 * (size_t)i);
    return 0;
}

// make_collision (nbody/src/lib.rs:504-528)
int gbh_ic_reference_collision(uint32_t n, uint64_t seed, double* bodies) {
    if (n < 4 || n > 1000000) {
        gbh::set_error("collision particle count must be in 4..=1000000");
        return -1;
    }
    uint32_t left = n / 2, right = n - left;
    double total = 5.0e10, half = total * 0.5;
    XorShift64 a(seed), b(seed + 1);
    for (uint32_t i = 0; i < left; ++i)
        reference_disc_body(a, half / (double)left, half, 3.0, -7.5, -1.2, 0.030, 0.010, 1.0, 0.86,
                            kDefaultSofteningKpc, bodies + 5 * (size_t)i);
    for (uint32_t i = 0; i < right; ++i)
        reference_disc_body(b, half / (double)right, half, 2.7, 7.5, 1.2, -0.030, -0.010, -1.0,
                            0.86, kDefaultSofteningKpc, bodies + 5 * (size_t)(left + i));
    return 0;
}

}  // extern "C"
