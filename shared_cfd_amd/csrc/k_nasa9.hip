// k_nasa9.hip -- set_thermo_prop of the cold-flow model
// (store/therm_lib/NASA/cold/thermal_model.f90:1-26): one lane per cell,
// cold_flow of shcfd_nasa9.h on (w(5:6), e_i), T started from p/(rho R) of
// the previous state.
#include "shcfd_device.h"
#include "shcfd_kernels.h"
#include "shcfd_nasa9.h"

namespace shcfd {

__global__ void __launch_bounds__(256) k_thermo_cold(DevCtx d, const Tile* tiles, const shcfd_nasa9* __restrict__ m) {
    using D = Dims<2>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    const long long c = cidx(B, t.i, t.j);
    const double rho = FLD(d, SHCFD_F_W, IW_RHO, c), u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
    const double qe = FLD(d, SHCFD_F_Q, 4, c);
    const double ei = (qe / rho) - (0.5 * ((u * u) + (v * v)));
    double T = FLD(d, SHCFD_F_W, IW_P, c) / (rho * FLD(d, SHCFD_F_W, D::ir, c));
    nasa9::ColdOut o;
    nasa9::cold_flow(m, FLD(d, SHCFD_F_W, IW_Y(0), c), FLD(d, SHCFD_F_W, IW_Y(1), c), ei, T, o);
    const double p = (rho * (nasa9::R_UNI / o.mwave)) * T;
    FLD(d, SHCFD_F_W, IW_P, c) = p;
    FLD(d, SHCFD_F_W, D::ig, c) = o.kappa;
    FLD(d, SHCFD_F_W, D::ih, c) = (qe + p) / rho;
    FLD(d, SHCFD_F_W, D::imu, c) = o.mu;
    FLD(d, SHCFD_F_W, D::ir, c) = nasa9::R_UNI / o.mwave;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        FLD(d, SHCFD_F_DHI, k, c) = o.dhi[k];
        FLD(d, SHCFD_F_VHI, k, c) = o.vhi[k];
    }
}

}  // namespace shcfd
