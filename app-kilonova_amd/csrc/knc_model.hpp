// knc_model.hpp -- the reference's initial models evaluated on the device.
//
// primitive_at / scalar_at of the InitialModel implementations in
// src/models/*.rs, for the ghost rows of model boundaries
// (BlockState::from_model on the boundary blocks at every stage's input
// time, src/scheme.rs:46-52) and for blocks born by excision
// (src/scheme.rs:197-205).  Each model reads its parameters from
// ModelDev::p in the field order of its Rust struct.  Expressions keep the
// reference's operation order; powf/exp/cos come from the device math
// library, so values can differ from a host evaluation in the last bit.
// A model that the reference would panic in (wind_shock outside its flares
// without a table, a table sampled out of range, halo_kilonova below the
// galactic midplane) reports failure instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knc_zone.hpp"

namespace knc {

enum {
    MODEL_KINETIC_BOMB = 1,
    MODEL_JET_IN_CLOUD = 2,
    MODEL_JET_IN_STAR = 3,
    MODEL_HALO_KILONOVA = 4,
    MODEL_WIND_SHOCK = 5,
};
constexpr int MODEL_NPARAM = 12;
constexpr double MODEL_LIGHT_SPEED = 3e10;   // src/physics/mod.rs:7

struct ModelDev {
    int32_t kind;
    int32_t table_rows;      // wind_shock: rows (r, u, d, h) of the initial data table
    double p[MODEL_NPARAM];
    const double* table;     // device copy, table_rows x 4
};

// ---- kinetic_bomb (src/models/kinetic_bomb.rs) and the shell of
// halo_kilonova (src/models/halo_kilonova.rs): p = external_medium_density,
// launch_radius, shell_thickness, kinetic_energy, shell_mass.
struct Shell {
    double v, dur, r_in, r_out;
};

__host__ __device__ inline Shell shell_at(double launch_radius, double thickness, double energy, double mass,
                                          double t)
{
    Shell s;
    s.v = sqrt(2.0 * energy / mass);
    s.dur = thickness / s.v;
    s.r_out = launch_radius + s.v * t;
    s.r_in = launch_radius + s.v * (t - s.dur);
    return s;
}

// Range::contains: r_in <= r < r_out.
__host__ __device__ inline bool in_shell(const Shell& s, double r) { return s.r_in <= r && r < s.r_out; }

__device__ inline Prim shell_primitive(const Shell& s, double mass, double r, double temperature)
{
    const double mdot = mass / s.dur;
    const double d = mdot / (4.0 * KNC_PI * r * r * s.v);
    return Prim{d, s.v / MODEL_LIGHT_SPEED, 0.0, d * temperature};
}

__device__ inline bool kinetic_bomb_at(const double* p, double r, double t, Prim* P, double* X)
{
    const Shell s = shell_at(p[1], p[2], p[3], p[4], t);
    if (in_shell(s, r)) {
        *P = shell_primitive(s, p[4], r, 1e-3);
        *X = 1.0;
    } else {
        const double x = r / p[1];
        const double d = p[0] * (x * x);
        *P = Prim{d, 0.0, 0.0, d * 1e-3};
        *X = 0.0;
    }
    return true;
}

// ---- jet_in_cloud (src/models/jet_in_cloud.rs): p = cloud_mass,
// engine_delay, engine_duration, engine_strength, engine_theta, engine_u,
// envelop_m1, envelop_slowest_beta, envelop_fastest_beta, envelop_psi.
__device__ inline bool jet_in_cloud_at(const double* p, double r, double q, double t, Prim* P, double* X)
{
    const double c = MODEL_LIGHT_SPEED;
    const double engine_u = p[5];
    const double engine_beta = engine_u / sqrt(1.0 + engine_u * engine_u);
    const double v_min = p[7] * c;
    const double v_jet = engine_beta * c;
    const double r_interface = v_min * t;
    const double r_head = v_jet * (t - p[1]);
    const double r_tail = v_jet * (t - p[1] - p[2]);
    const bool nozzle = q < p[4] || q > KNC_PI - p[4];
    double f, u;
    if (nozzle && r < r_head && r > r_tail) {
        const double engine_gamma = sqrt(1.0 + engine_u * engine_u);
        const double e = p[3] * p[0];
        const double l = e / (4.0 * KNC_PI * p[2]);
        f = l / engine_gamma;
        u = engine_u;
        *X = 1e2;
    } else if (r > r_interface) {
        const double b = fmin(r / t / c, p[8]);
        u = b / sqrt(1.0 - b * b);
        const double s = b;
        const double g = pow(s, -1.0 / p[9]) * pow(1.0 - s * s, 0.5 / p[9] - 1.0);
        f = p[6] / (4.0 * KNC_PI * p[9] * t) * g;
        *X = 1e-2;
    } else {
        const double b = p[7];
        u = b / sqrt(1.0 - b * b);
        f = p[0] / (4.0 * KNC_PI * p[1]);
        *X = 1e0;
    }
    const double d = f / (r * r * u) / c;
    *P = Prim{d, u, 0.0, d * 1e-3};
    return true;
}

// ---- jet_in_star (src/models/jet_in_star.rs): p = star_mass,
// engine_duration, engine_energy, engine_theta, engine_u, envelope_radius,
// envelope_mass, volume_factor.  Constants of Duffell & MacFadyen (2015).
namespace jis {
constexpr double R0 = 7e10;
constexpr double M0 = 2e33;
constexpr double RHO_C = 3e7 * M0 / (1.33 * KNC_PI * R0 * R0 * R0);
constexpr double R1 = 0.0017 * R0;
constexpr double R2 = 0.0125 * R0;
constexpr double R3 = 0.65 * R0;
constexpr double K1 = 3.24;
constexpr double K2 = 2.57;
constexpr double N = 16.7;
constexpr double RHO_WIND = 1e-9 * M0 / (1.33 * KNC_PI * R0 * R0 * R0);
constexpr double R_NOZZ = 0.01 * R0;
constexpr double R_ENV = 1.2 * R0;
constexpr double ALPHA = 2.5;
}  // namespace jis

__device__ inline bool jet_in_star_at(const double* p, double r, double q, double t, Prim* P, double* X)
{
    using namespace jis;
    const double c = MODEL_LIGHT_SPEED;
    const double engine_u = p[4];
    const double engine_beta = engine_u / sqrt(1.0 + engine_u * engine_u);
    const double r_head = engine_beta * c * t;
    const bool nozzle = q < p[3] || q > KNC_PI - p[3];
    const double er = p[5];
    const double rho_env = p[6] / (4.0 * KNC_PI * (er * er) * (er - R3) * p[7]);
    double d, u = 0.0;
    if (nozzle && r < r_head) {
        // nozzle function (Duffell & MacFadyen 2015) and jet mass rate
        const double r0 = R_NOZZ / R0;
        const double q2 = p[3] * p[3];
        const double n0 = 4.0 * KNC_PI * r0 * r0 * r0 * (1.0 - exp(-2.0 / q2)) * q2;
        const double g = (r / R_NOZZ) * exp(-pow(r / R_NOZZ, 2.0) / 2.0) * exp((pow(cos(q), 2.0) - 1.0) / q2);
        const double engine_gamma = sqrt(1.0 + engine_u * engine_u);
        const double l = g / n0 * p[2] / (4.0 * KNC_PI * p[1]);
        const double rate = l / (engine_gamma * c * c);
        d = rate / (r * r * engine_u * c);
        u = engine_u;
        *X = 1e2;
    } else if (r < R3) {
        const double num = RHO_C * pow(1.0 - r / R3, N);
        const double denom = 1.0 + pow(r / R1, K1) / (1.0 + pow(r / R2, K2));
        d = num / denom + rho_env * pow(r / R3, -2.0);
        *X = 1e0;
    } else if (R3 < r && r < er) {
        d = rho_env * pow(r / R3, -ALPHA);
        *X = 1e-2 * pow(r / R3, -2.0);
    } else {
        d = RHO_WIND * pow(r / R_ENV, -2.0);
        *X = 1e-5 * pow(r / R_ENV, -2.0);
    }
    *P = Prim{d, u, 0.0, d * 1e-10};
    return true;
}

// ---- halo_kilonova (src/models/halo_kilonova.rs, src/galmod.rs):
// p = altitude, launch_radius, shell_thickness, kinetic_energy, shell_mass,
// radial_distance.  Outside the shell the gas is the thin disk of a
// Miyamoto-Nagai galactic model in hydrostatic balance: the density at
// (radial_distance, z) and the pressure from p0 = 1e-3 at z = 0 integrated
// up to z in 1000 RK4 steps of the vertical gravity.
namespace gal {
constexpr double G = 6.67e-8;
constexpr double M_B = 3.377e43;
constexpr double A_B = 8.98e20;
constexpr double V_H = 1.923e7;
constexpr double A_H = 9.26e22;
constexpr double M_S = 1.538e44;
constexpr double A_S = 1.461e22;
constexpr double B_S = 1.790e21;
constexpr double A_G = 1.461e22;
constexpr double B_G = 7.035e23;

// x^1.5 and x^2.5 of a positive x, the reference's powf(1.5) and
// powf(2.5) of the galactic model, as x sqrt(x) and (x x) sqrt(x): within
// two roundings of the power, like the device pow, at a tenth of its
// instructions.  The ghost rows evaluate them some 20 000 times per zone.
__device__ inline double pow_1p5(double x) { return x * sqrt_rn(x); }
__device__ inline double pow_2p5(double x) { return (x * x) * sqrt_rn(x); }

// GalacticModel::density(r, z).thin_disk.  powi(n) of the reference is
// repeated multiplication (x*x, x*(x*x), (x*x)*(x*x)).
__device__ inline double thin_disk_density(double r, double z)
{
    const double zz = z * z + B_S * B_S;
    const double zb = sqrt(zz);
    const double zb3 = zz * zb;   // pow(zz, 1.5), with zb = sqrt(zz)
    const double z2 = z * z, b2 = B_S * B_S;
    const double z4 = z2 * z2, b4 = b2 * b2, a3 = A_S * (A_S * A_S);
    const double za = zb + A_S;
    const double dz = -M_S / 4.0 / KNC_PI *
                      ((4.0 * A_S * z4 + zb * ((2.0 * A_S * A_S - r * r) * z * z - B_S * B_S * r * r -
                                               3.0 * A_S * A_S * B_S * B_S) +
                        zb3 * (2.0 * z * z - B_S * B_S) + A_S * B_S * B_S * z * z - A_S * B_S * B_S * r * r -
                        3.0 * A_S * b4 - a3 * B_S * B_S) /
                       (zb3 * pow_2p5(za * za + r * r)));
    const double as = A_S + zb;
    const double w = r * r + as * as;
    const double dr = M_S / 4.0 / KNC_PI * ((2.0 * w - 3.0 * r * r) / pow_2p5(w));
    return dz + dr;
}

// GalacticModel::g_field_z(r, z).total() (its thin and thick disk terms use the
// bulge mass m_b, as the reference does)
__device__ inline double g_field_z(double r, double z)
{
    const double bulge = -G * M_B * z / pow_1p5(r * r + z * z + A_B * A_B);
    const double zs = sqrt(z * z + B_S * B_S);
    const double as = A_S + zs;
    const double thin = -G * M_B * z * (A_S + zs) / pow_1p5(r * r + as * as) / zs;
    const double zg = sqrt(z * z + B_G * B_G);
    const double ag = A_G + zg;
    const double thick = -G * M_B * z * (A_G + zg) / pow_1p5(r * r + ag * ag) / zg;
    const double halo = -V_H * V_H * z / (r * r + z * z + A_H * A_H);
    return bulge + thin + thick + halo;
}

__device__ inline double pressure_difference_rk4(double r, double z, double dz)
{
    const double k1 = thin_disk_density(r, z) * g_field_z(r, z);
    const double k2 = thin_disk_density(r, z + 0.5 * dz) * g_field_z(r, z + 0.5 * dz);
    const double k3 = k2;
    const double k4 = thin_disk_density(r, z + dz) * g_field_z(r, z + dz);
    return dz * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
}
}  // namespace gal

// The disk part of a halo_kilonova zone: the altitude z, the step dz of the
// pressure integration and the density.
struct HaloDisk {
    double z, dz, d;
};

// At most this many RK4 steps of the pressure integration (1000 up to z,
// and one more where the accumulated altitude rounds below z).
constexpr int HALO_MAX_STEPS = 2000;

// halo_kilonova up to the pressure integration: 0 for a zone in the shell
// (*P and *X set), -1 below the midplane (the reference panics), 1 for the
// disk (*h set; the pressure is 1e-3 plus the RK4 steps at the altitudes
// 0, dz, 2 dz, ... accumulated in that order while below z).
__device__ inline int halo_kilonova_begin(const double* p, double r, double q, double t, Prim* P, double* X,
                                          HaloDisk* h)
{
    const double z = r * cos(q) + p[0];
    const Shell s = shell_at(p[1], p[2], p[3], p[4], t);
    if (in_shell(s, r)) {
        *P = shell_primitive(s, p[4], r, 1e-3);
        *X = 1.0;
        return 0;
    }
    if (!(z > 0.0)) return -1;
    h->z = z;
    h->d = gal::thin_disk_density(p[5], z);
    // vertical_pressure_profile(r, z, z * 1e-3, 1e-3).last()
    h->dz = z * 1e-3;
    return 1;
}

__device__ inline bool halo_kilonova_at(const double* p, double r, double q, double t, Prim* P, double* X)
{
    HaloDisk h;
    const int c = halo_kilonova_begin(p, r, q, t, P, X, &h);
    if (c <= 0) return c == 0;
    double zz = 0.0, pr = 1e-3;
    for (int n = 0; zz < h.z && n < HALO_MAX_STEPS; ++n) {
        pr += gal::pressure_difference_rk4(r, zz, h.dz);
        zz += h.dz;
    }
    *P = Prim{h.d, 0.0, 0.0, pr};
    *X = 0.0;
    return true;
}

// ---- wind_shock (src/models/wind_shock.rs): p = wind_mass_outflow_rate,
// wind_gamma_beta, wind_pressure, post_shock_pressure, shock_location,
// post_shock_gamma_beta, flare_outflow_rate, flare_gamma_beta, flare_time,
// flare_duration; three flares at flare_time + {0, 3, 6}, otherwise the
// initial data table (LookupTable<4>::sample, src/lookup_table_v2.rs).

// LookupTable::sample at x of a table with strictly increasing column 0:
// linear interpolation between rows (i-1, i), i the first row with x_i >= x;
// false (the reference panics) unless x_0 < x <= x_last.
__host__ __device__ inline bool table_sample(const double* rows, int n, double x, double out[4])
{
    if (!(x > rows[0] && x <= rows[4 * (n - 1)])) return false;
    int lo = 0, hi = n - 1;   // rows[lo] < x <= rows[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (rows[4 * mid] < x) lo = mid; else hi = mid;
    }
    const double* a = rows + 4 * (hi - 1);
    const double* b = rows + 4 * hi;
    for (int k = 0; k < 4; ++k) out[k] = a[k] + (x - a[0]) * (b[k] - a[k]) / (b[0] - a[0]);
    return true;
}

__device__ inline bool wind_shock_at(const ModelDev& m, double r, double t, Prim* P, double* X)
{
    const double* p = m.p;
    const double c = MODEL_LIGHT_SPEED;
    *X = 0.0;
    const double ft = p[8], fd = p[9];
    for (int k = 0; k < 3; ++k) {
        const double t0 = k == 0 ? ft : ft + 3.0 * k;
        if (t >= t0 && t < t0 + fd) {
            const double u = p[7];
            const double n = p[6] / (4.0 * KNC_PI * r * r * u * c);
            const double rho = n * (t0 + fd - t) / fd;
            *P = Prim{rho, u, 0.0, rho * 1e-6};
            return true;
        }
    }
    if (m.table_rows < 2) return false;
    double s[4];
    if (!table_sample(m.table, m.table_rows, r, s)) return false;
    const double u = s[1], d = s[2], h = s[3];
    const double mu = h / c / c - 1.0;
    const double e = mu / (4.0 / 3.0);
    *P = Prim{d, u, 0.0, d * e * (4.0 / 3.0 - 1.0)};
    return true;
}

// primitive_at and scalar_at of the model at (r, q) and time t (the
// AnyPrimitive fields map one to one onto Prim, as `interpret` does,
// src/physics/relativistic_hydro.rs:110-112).
__device__ inline bool model_at(const ModelDev& m, double r, double q, double t, Prim* P, double* X)
{
    switch (m.kind) {
    case MODEL_KINETIC_BOMB: return kinetic_bomb_at(m.p, r, t, P, X);
    case MODEL_JET_IN_CLOUD: return jet_in_cloud_at(m.p, r, q, t, P, X);
    case MODEL_JET_IN_STAR: return jet_in_star_at(m.p, r, q, t, P, X);
    case MODEL_HALO_KILONOVA: return halo_kilonova_at(m.p, r, q, t, P, X);
    case MODEL_WIND_SHOCK: return wind_shock_at(m, r, t, P, X);
    default: return false;
    }
}

}  // namespace knc
