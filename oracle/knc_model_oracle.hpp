// knc_model_oracle.hpp -- CPU restatement of the reference's initial models
// and block mesh (test infrastructure; see knc_model_oracle.cpp).
#pragma once

namespace om {

// Same layouts as knc_model / knc_mesh of the library's C ABI.
struct Model {
    int kind;          // 1 kinetic_bomb, 2 jet_in_cloud, 3 jet_in_star, 4 halo_kilonova, 5 wind_shock
    int table_rows;
    double params[12];
    const double* table;
};

struct Mesh {
    double reference_radius;
    int num_radial_zones;   // per decade, 0 = None
    int num_polar_zones;
    int block_size;
    int pad;
};

// primitive_at (rho, u_r, u_theta, p) and scalar_at; false where the
// reference panics.
bool primitive_at(const Model& m, double r, double q, double t, double P[4], double* X);
void block_vertices(const Mesh& mesh, int index, double* rv, double* qv);

}  // namespace om
