// knc_model_oracle.cpp -- CPU restatement of the reference's initial models
// and of the radial block mesh.  Test infrastructure: only the tests, the
// smoke check and bench.py's CPU baseline load it, as the checker of the
// device models (app-kilonova_amd/csrc/knc_model.hpp).
//
// Follows src/models/{kinetic_bomb,jet_in_cloud,jet_in_star,halo_kilonova,
// wind_shock}.rs, src/galmod.rs, src/lookup_table_v2.rs and the mesh
// geometry of src/mesh.rs:202-264,341-375, in the reference's operation
// order, with the host's libm for powf/exp/cos/ln.
#include <cmath>
#include <cstdint>
#include <vector>

#include "knc_model_oracle.hpp"

namespace om {

namespace {

const double PI = 3.14159265358979323846;
const double C = 3e10;   // LIGHT_SPEED, src/physics/mod.rs:7

// Rust's powi(n) for the small constant n the models use.
double sq(double x) { return x * x; }

// kinetic_bomb.rs:35-55 / halo_kilonova.rs:30-44
struct ShellExtent {
    double velocity, duration, lower, upper;
};

ShellExtent shell(double launch_radius, double thickness, double kinetic_energy, double mass, double t)
{
    ShellExtent s;
    s.velocity = std::sqrt(2.0 * kinetic_energy / mass);
    s.duration = thickness / s.velocity;
    s.upper = launch_radius + s.velocity * t;
    s.lower = launch_radius + s.velocity * (t - s.duration);
    return s;
}

bool contains(const ShellExtent& s, double r) { return r >= s.lower && r < s.upper; }

void shell_state(const ShellExtent& s, double mass, double r, double temperature, double P[4])
{
    const double mdot = mass / s.duration;
    const double d = mdot / (4.0 * PI * r * r * s.velocity);
    P[0] = d;
    P[1] = s.velocity / C;
    P[2] = 0.0;
    P[3] = d * temperature;
}

// kinetic_bomb.rs:79-112
bool kinetic_bomb(const double* m, double r, double t, double P[4], double* X)
{
    const ShellExtent s = shell(m[1], m[2], m[3], m[4], t);
    if (contains(s, r)) {
        shell_state(s, m[4], r, 1e-3, P);
        *X = 1.0;
        return true;
    }
    const double d = m[0] * sq(r / m[1]);
    P[0] = d;
    P[1] = 0.0;
    P[2] = 0.0;
    P[3] = d * 1e-3;
    *X = 0.0;
    return true;
}

// jet_in_cloud.rs:60-258
enum JicZone { CLOUD, ENVELOPE, JET };

bool jet_in_cloud(const double* m, double r, double q, double t, double P[4], double* X)
{
    const double cloud_mass = m[0], delay = m[1], duration = m[2], strength = m[3], theta = m[4];
    const double eu = m[5], m1 = m[6], slowest = m[7], fastest = m[8], psi = m[9];
    const double beta_jet = eu / std::sqrt(1.0 + sq(eu));
    const double v_jet = beta_jet * C;
    JicZone z;
    if ((q < theta || q > PI - theta) && r < v_jet * (t - delay) && r > v_jet * (t - delay - duration))
        z = JET;
    else if (r > slowest * C * t)
        z = ENVELOPE;
    else
        z = CLOUD;
    double u, f;
    switch (z) {
    case CLOUD:
        u = slowest / std::sqrt(1.0 - slowest * slowest);
        f = cloud_mass / (4.0 * PI * delay);
        *X = 1e+0;
        break;
    case ENVELOPE: {
        const double b = std::fmin(r / t / C, fastest);
        u = b / std::sqrt(1.0 - b * b);
        const double g = std::pow(b, -1.0 / psi) * std::pow(1.0 - b * b, 0.5 / psi - 1.0);
        f = m1 / (4.0 * PI * psi * t) * g;
        *X = 1e-2;
        break;
    }
    default: {
        const double gamma = std::sqrt(1.0 + eu * eu);
        const double e = strength * cloud_mass;
        f = e / (4.0 * PI * duration) / gamma;
        u = eu;
        *X = 1e+2;
    }
    }
    const double d = f / (r * r * u) / C;
    P[0] = d;
    P[1] = u;
    P[2] = 0.0;
    P[3] = d * 1e-3;
    return true;
}

// jet_in_star.rs:1-237 (constants of Duffell & MacFadyen 2015)
const double JR0 = 7e10;
const double JM0 = 2e33;
const double RHO_C = 3e7 * JM0 / (1.33 * PI * JR0 * JR0 * JR0);
const double JR1 = 0.0017 * JR0;
const double JR2 = 0.0125 * JR0;
const double JR3 = 0.65 * JR0;
const double RHO_WIND = 1e-9 * JM0 / (1.33 * PI * JR0 * JR0 * JR0);
const double R_NOZZ = 0.01 * JR0;
const double R_ENV = 1.2 * JR0;

bool jet_in_star(const double* m, double r, double q, double t, double P[4], double* X)
{
    const double duration = m[1], energy = m[2], theta = m[3], eu = m[4];
    const double env_r = m[5], env_m = m[6], vf = m[7];
    const double beta = eu / std::sqrt(1.0 + sq(eu));
    double d, u = 0.0;
    const double rho_env = env_m / (4.0 * PI * sq(env_r) * (env_r - JR3) * vf);
    if ((q < theta || q > PI - theta) && r < beta * C * t) {
        const double r0 = R_NOZZ / JR0;
        const double q2 = sq(theta);
        const double n0 = 4.0 * PI * r0 * r0 * r0 * (1.0 - std::exp(-2.0 / q2)) * q2;
        const double g = (r / R_NOZZ) * std::exp(-std::pow(r / R_NOZZ, 2.0) / 2.0) *
                         std::exp((std::pow(std::cos(q), 2.0) - 1.0) / q2);
        const double nozzle = g / n0;
        const double l = nozzle * energy / (4.0 * PI * duration);
        const double rate = l / (std::sqrt(1.0 + eu * eu) * C * C);
        d = rate / (r * r * eu * C);
        u = eu;
        *X = 1e+2;
    } else if (r < JR3) {
        const double core = RHO_C * std::pow(1.0 - r / JR3, 16.7) /
                            (1.0 + std::pow(r / JR1, 3.24) / (1.0 + std::pow(r / JR2, 2.57)));
        d = core + rho_env * std::pow(r / JR3, -2.0);
        *X = 1e+0;
    } else if (JR3 < r && r < env_r) {
        d = rho_env * std::pow(r / JR3, -2.5);
        *X = 1e-2 * std::pow(r / JR3, -2.0);
    } else {
        d = RHO_WIND * std::pow(r / R_ENV, -2.0);
        *X = 1e-5 * std::pow(r / R_ENV, -2.0);
    }
    P[0] = d;
    P[1] = u;
    P[2] = 0.0;
    P[3] = d * 1e-10;
    return true;
}

// galmod.rs: the thin-disk density, the vertical field and the RK4
// hydrostatic pressure profile
struct Galaxy {
    double g = 6.67e-8, m_b = 3.377e43, a_b = 8.98e20, v_h = 1.923e7, a_h = 9.26e22, m_s = 1.538e44,
           a_s = 1.461e22, b_s = 1.790e21, a_g = 1.461e22, b_g = 7.035e23;

    double thin_disk_density(double r, double z) const
    {
        const double zs = z * z + b_s * b_s;
        const double rz = -m_s / 4.0 / PI *
                          ((4.0 * a_s * sq(sq(z)) +
                            std::sqrt(zs) * ((2.0 * a_s * a_s - r * r) * z * z - b_s * b_s * r * r -
                                             3.0 * a_s * a_s * b_s * b_s) +
                            std::pow(zs, 1.5) * (2.0 * z * z - b_s * b_s) + a_s * b_s * b_s * z * z -
                            a_s * b_s * b_s * r * r - 3.0 * a_s * sq(sq(b_s)) - a_s * sq(a_s) * b_s * b_s) /
                           (std::pow(zs, 1.5) * std::pow(sq(std::sqrt(zs) + a_s) + r * r, 2.5)));
        const double rr = m_s / 4.0 / PI *
                          ((2.0 * (r * r + sq(a_s + std::sqrt(zs))) - 3.0 * r * r) /
                           std::pow(r * r + sq(a_s + std::sqrt(zs)), 2.5));
        return rz + rr;
    }

    double field_z(double r, double z) const
    {
        const double bulge = -g * m_b * z / std::pow(r * r + z * z + a_b * a_b, 1.5);
        const double ts = std::sqrt(z * z + b_s * b_s);
        const double thin = -g * m_b * z * (a_s + ts) / std::pow(r * r + sq(a_s + ts), 1.5) / ts;
        const double tg = std::sqrt(z * z + b_g * b_g);
        const double thick = -g * m_b * z * (a_g + tg) / std::pow(r * r + sq(a_g + tg), 3.0 / 2.0) / tg;
        const double halo = -v_h * v_h * z / (r * r + z * z + a_h * a_h);
        return bulge + thin + thick + halo;
    }

    double rk4(double r, double z, double dz) const
    {
        const double k1 = thin_disk_density(r, z) * field_z(r, z);
        const double k2 = thin_disk_density(r, z + 0.5 * dz) * field_z(r, z + 0.5 * dz);
        const double k3 = thin_disk_density(r, z + 0.5 * dz) * field_z(r, z + 0.5 * dz);
        const double k4 = thin_disk_density(r, z + dz) * field_z(r, z + dz);
        return dz * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
    }
};

bool halo_kilonova(const double* m, double r, double q, double t, double P[4], double* X)
{
    const double z = r * std::cos(q) + m[0];
    const ShellExtent s = shell(m[1], m[2], m[3], m[4], t);
    if (contains(s, r)) {
        shell_state(s, m[4], r, 1e-3, P);
        *X = 1.0;
        return true;
    }
    *X = 0.0;
    if (!(z > 0.0)) return false;
    const Galaxy gal;
    const double dz = z * 1e-3;
    double zz = 0.0, p = 1e-3;
    while (zz < z) {
        p += gal.rk4(r, zz, dz);
        zz += dz;
    }
    P[0] = gal.thin_disk_density(m[5], z);
    P[1] = 0.0;
    P[2] = 0.0;
    P[3] = p;
    return true;
}

// lookup_table_v2.rs:104-135: linear interpolation between the rows
// straddling x, found by binary search; out of range -> the reference panics
bool sample(const double* rows, int n, double x, double out[4])
{
    if (!(x > rows[0])) return false;
    if (!(x <= rows[4 * (n - 1)])) return false;
    int lo = 0, hi = n;   // first index with row >= x, in [1, n-1]
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (rows[4 * mid] < x) lo = mid + 1; else hi = mid;
    }
    const int i1 = lo, i0 = lo - 1;
    for (int k = 0; k < 4; ++k) {
        const double x0 = rows[4 * i0], y0 = rows[4 * i0 + k], x1 = rows[4 * i1], y1 = rows[4 * i1 + k];
        out[k] = y0 + (x - x0) * (y1 - y0) / (x1 - x0);
    }
    return true;
}

// wind_shock.rs:79-150
bool wind_shock(const Model& md, double r, double t, double P[4], double* X)
{
    const double* m = md.params;
    *X = 0.0;
    const double starts[3] = {m[8], m[8] + 3.0, m[8] + 6.0};
    for (double t0 : starts) {
        if (t >= t0 && t < t0 + m[9]) {
            const double u = m[7];
            const double n = m[6] / (4.0 * PI * r * r * u * C);
            const double rho = n * (t0 + m[9] - t) / m[9];
            P[0] = rho;
            P[1] = u;
            P[2] = 0.0;
            P[3] = rho * 1e-6;
            return true;
        }
    }
    if (md.table_rows < 2) return false;
    double s[4];
    if (!sample(md.table, md.table_rows, r, s)) return false;
    const double mu = s[3] / C / C - 1.0;
    const double e = mu / (4.0 / 3.0);
    P[0] = s[2];
    P[1] = s[1];
    P[2] = 0.0;
    P[3] = s[2] * e * (4.0 / 3.0 - 1.0);
    return true;
}

}  // namespace

bool primitive_at(const Model& m, double r, double q, double t, double P[4], double* X)
{
    switch (m.kind) {
    case 1: return kinetic_bomb(m.params, r, t, P, X);
    case 2: return jet_in_cloud(m.params, r, q, t, P, X);
    case 3: return jet_in_star(m.params, r, q, t, P, X);
    case 4: return halo_kilonova(m.params, r, q, t, P, X);
    case 5: return wind_shock(m, r, t, P, X);
    default: return false;
    }
}

// Mesh::subgrid_extent + SphericalPolarGrid::vertex_coordinate_signed.
void block_vertices(const Mesh& mesh, int index, double* rv, double* qv)
{
    const double zone = mesh.num_radial_zones > 0 ? 1.0 / mesh.num_radial_zones : PI / mesh.num_polar_zones;
    const double block = mesh.block_size * zone;
    const double inner = mesh.reference_radius * std::pow(1.0 + block, (double)index);
    const double outer = mesh.reference_radius * std::pow(1.0 + block, (double)index + 1.0);
    double q0 = 0.0, q1 = PI;
    if (mesh.num_polar_zones == 1) {
        q0 = PI * 0.5 - zone;
        q1 = PI * 0.5 + zone;
    }
    const double y0 = std::log(inner) / std::log(10.0);
    const double y1 = std::log(outer) / std::log(10.0);
    const double dy = (y1 - y0) / mesh.block_size;
    const double dq = (q1 - q0) / mesh.num_polar_zones;
    if (rv)
        for (int i = 0; i <= mesh.block_size; ++i) rv[i] = std::pow(10.0, y0 + dy * i);
    if (qv)
        for (int j = 0; j <= mesh.num_polar_zones; ++j) qv[j] = q0 + dq * j;
}

}  // namespace om

extern "C" {

// primitive_at / scalar_at at n points; returns the index of the first
// point the reference would panic at, or -1.
long long ko_model_eval(const om::Model* m, long long n, const double* r, const double* q, double t, double* prim,
                        double* scalar)
{
    long long bad = -1;
    for (long long k = 0; k < n; ++k) {
        double P[4] = {0.0, 0.0, 0.0, 0.0}, X = 0.0;
        if (!om::primitive_at(*m, r[k], q[k], t, P, &X) && bad < 0) bad = k;
        for (int c = 0; c < 4; ++c) prim[4 * k + c] = P[c];
        scalar[k] = X;
    }
    return bad;
}

void ko_mesh_block_vertices(const om::Mesh* mesh, int index, double* rv, double* qv)
{
    om::block_vertices(*mesh, index, rv, qv);
}

// mesh.subgrid_extent(index).outer_radius (src/mesh.rs:358-372), the radius
// add_remove_blocks compares with the excision surfaces (src/scheme.rs:190,195).
double ko_mesh_outer_radius(const om::Mesh* mesh, int index)
{
    const double zone = mesh->num_radial_zones > 0 ? 1.0 / mesh->num_radial_zones : om::PI / mesh->num_polar_zones;
    const double block = mesh->block_size * zone;
    return mesh->reference_radius * std::pow(1.0 + block, (double)index + 1.0);
}

}  // extern "C"
