// k_geom.hip -- grid metrics from node coordinates (init time):
// dim/2d/geometry.f90:79-169 (centres, Vol, dsi/dsj, vni/vnj),
// LU-SGS/2d/sch_lusgs.raw.f90:143-181 (dsci/dscj, vnci/vncj),
// viscosity/viscous/no-nV/2d/sch_viscous.f90:1-54 (geojaci/geojacj).
// Axisymmetric grids (dim/q2d/geometry.f90:107-141, LU-SGS/q2d/
// sch_lusgs.raw.f90:151-186, viscous/no-nV/q2d/sch_viscous.f90:1-29): every
// volume and face measure is per radian about the x axis -- Vol is the sum
// over the cell's two triangles of the triangle area times its mean vertex r
// (exact for the revolved triangle), dsi/dsj are face lengths times the mean
// r of their two nodes, dsci/dscj carry the mean r of the two mid-face points
// they join, Area keeps the planar cell area (weight of the pressure and hoop
// stress sources) and geojacc holds the cell-centred metrics of the hoop
// stress.
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

#define XH(i, j) FLD(d, SHCFD_F_XH, 0, cidx(B, i, j))
#define RH(i, j) FLD(d, SHCFD_F_RH, 0, cidx(B, i, j))
#define XC(i, j) FLD(d, SHCFD_F_X, 0, cidx(B, i, j))
#define RC(i, j) FLD(d, SHCFD_F_R, 0, cidx(B, i, j))

__device__ inline double area_triangle(double a1, double a2, double b1, double b2, double c1, double c2) {
    return 0.5 * (((a1 * (b2 - c2)) + (b1 * (c2 - a2))) + (c1 * (a2 - b2)));
}
__device__ inline double dis_point(double x1, double y1, double x2, double y2) {
    return sqrt(((x1 - x2) * (x1 - x2)) + ((y1 - y2) * (y1 - y2)));
}

// pass 0: interior centres, Vol, face lengths/normals, cell-centre metrics
__global__ void __launch_bounds__(256) k_geom_a(DevCtx d, const Tile* tiles) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j, ni = B.ni, nj = B.nj;
    const long long c = cidx(B, i, j);
    const bool ax = d.p.dimension == SHCFD_DIM_AXISYMMETRIC;
    if (i >= 1 && i <= ni && j >= 1 && j <= nj) {
        XC(i, j) = 0.25 * (((XH(i - 1, j - 1) + XH(i, j - 1)) + XH(i - 1, j)) + XH(i, j));
        RC(i, j) = 0.25 * (((RH(i - 1, j - 1) + RH(i, j - 1)) + RH(i - 1, j)) + RH(i, j));
        double aabd = area_triangle(XH(i - 1, j - 1), RH(i - 1, j - 1), XH(i, j - 1), RH(i, j - 1), XH(i - 1, j), RH(i - 1, j));
        double abcd = area_triangle(XH(i, j - 1), RH(i, j - 1), XH(i, j), RH(i, j), XH(i - 1, j), RH(i - 1, j));
        if (ax) {
            FLD(d, SHCFD_F_AREA, 0, c) = aabd + abcd;
            FLD(d, SHCFD_F_VOL, 0, c) = ((((RH(i - 1, j - 1) + RH(i, j - 1)) + RH(i - 1, j)) * aabd) / 3.0) +
                                        ((((RH(i, j - 1) + RH(i, j)) + RH(i - 1, j)) * abcd) / 3.0);
        } else {
            FLD(d, SHCFD_F_AREA, 0, c) = 0.0;
            FLD(d, SHCFD_F_VOL, 0, c) = aabd + abcd;
        }
        double a1 = (XH(i - 1, j - 1) + XH(i, j - 1)) * 0.5, a2 = (RH(i - 1, j - 1) + RH(i, j - 1)) * 0.5;
        double b1 = (XH(i, j - 1) + XH(i, j)) * 0.5, b2 = (RH(i, j - 1) + RH(i, j)) * 0.5;
        double c1 = (XH(i - 1, j) + XH(i, j)) * 0.5, c2 = (RH(i - 1, j) + RH(i, j)) * 0.5;
        double d1 = (XH(i - 1, j - 1) + XH(i - 1, j)) * 0.5, d2 = (RH(i - 1, j - 1) + RH(i - 1, j)) * 0.5;
        double dsci = sqrt(((c1 - a1) * (c1 - a1)) + ((c2 - a2) * (c2 - a2)));
        double dscj = sqrt(((b1 - d1) * (b1 - d1)) + ((b2 - d2) * (b2 - d2)));
        FLD(d, SHCFD_F_VNCI, 0, c) = (b1 - d1) / dscj;
        FLD(d, SHCFD_F_VNCI, 1, c) = (b2 - d2) / dscj;
        FLD(d, SHCFD_F_VNCJ, 0, c) = (c1 - a1) / dsci;
        FLD(d, SHCFD_F_VNCJ, 1, c) = (c2 - a2) / dsci;
        if (ax) {
            dsci = (dsci * (c2 + a2)) * 0.5;
            dscj = (dscj * (b2 + d2)) * 0.5;
        }
        FLD(d, SHCFD_F_DSCI, 0, c) = dsci;
        FLD(d, SHCFD_F_DSCJ, 0, c) = dscj;
    }
    if (i >= 1 && i <= ni && j <= nj) {  // j-face (i, j), j in [0, nj]
        double ds = dis_point(XH(i - 1, j), RH(i - 1, j), XH(i, j), RH(i, j));
        FLD(d, SHCFD_F_DSJ, 0, c) = ax ? (ds * (RH(i - 1, j) + RH(i, j))) / 2.0 : ds;
        double nx = -(RH(i, j) - RH(i - 1, j));
        double nr = XH(i, j) - XH(i - 1, j);
        double sd = sqrt((nx * nx) + (nr * nr));
        FLD(d, SHCFD_F_VNJ, 0, c) = nx / sd;
        FLD(d, SHCFD_F_VNJ, 1, c) = nr / sd;
    }
    if (i <= ni && j >= 1 && j <= nj) {  // i-face (i, j), i in [0, ni]
        double ds = dis_point(XH(i, j - 1), RH(i, j - 1), XH(i, j), RH(i, j));
        FLD(d, SHCFD_F_DSI, 0, c) = ax ? (ds * (RH(i, j - 1) + RH(i, j))) / 2.0 : ds;
        double nx = RH(i, j) - RH(i, j - 1);
        double nr = -(XH(i, j) - XH(i, j - 1));
        double sd = sqrt((nx * nx) + (nr * nr));
        FLD(d, SHCFD_F_VNI, 0, c) = nx / sd;
        FLD(d, SHCFD_F_VNI, 1, c) = nr / sd;
    }
}

// pass 1: ghost centres on i-sides (j in 1..nj); pass 2: j-sides (i in 0..ni+1)
__global__ void k_geom_ghost(DevCtx d, int pass) {
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    const DevBlock B = d.blk[b];
    const int ni = B.ni, nj = B.nj;
    if (pass == 1) {
        int j = gid + 1;
        if (j > nj) return;
        XC(0, j) = 2.0 * XC(1, j) - XC(2, j);
        RC(0, j) = 2.0 * RC(1, j) - RC(2, j);
        XC(ni + 1, j) = 2.0 * XC(ni, j) - XC(ni - 1, j);
        RC(ni + 1, j) = 2.0 * RC(ni, j) - RC(ni - 1, j);
    } else {
        int i = gid;
        if (i > ni + 1) return;
        XC(i, 0) = 2.0 * XC(i, 1) - XC(i, 2);
        RC(i, 0) = 2.0 * RC(i, 1) - RC(i, 2);
        XC(i, nj + 1) = 2.0 * XC(i, nj) - XC(i, nj - 1);
        RC(i, nj + 1) = 2.0 * RC(i, nj) - RC(i, nj - 1);
    }
}

__global__ void __launch_bounds__(256) k_geojac(DevCtx d, const Tile* tiles) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j, ni = B.ni, nj = B.nj;
    const long long c = cidx(B, i, j);
    if (i <= ni && j >= 1 && j <= nj) {
        long long st[6];
        viscous_stencil<0>(c, B.pitch, st);
        double xc[6], rc[6], g[4];
        for (int s = 0; s < 6; s++) {
            xc[s] = FLD(d, SHCFD_F_X, 0, st[s]);
            rc[s] = FLD(d, SHCFD_F_R, 0, st[s]);
        }
        geojac_t<0, double>(xc, rc, g);
        for (int k = 0; k < 4; k++) FLD(d, SHCFD_F_GEOJACI, k, c) = g[k];
    }
    if (i >= 1 && i <= ni && j <= nj) {
        long long st[6];
        viscous_stencil<1>(c, B.pitch, st);
        double xc[6], rc[6], g[4];
        for (int s = 0; s < 6; s++) {
            xc[s] = FLD(d, SHCFD_F_X, 0, st[s]);
            rc[s] = FLD(d, SHCFD_F_R, 0, st[s]);
        }
        geojac_t<1, double>(xc, rc, g);
        for (int k = 0; k < 4; k++) FLD(d, SHCFD_F_GEOJACJ, k, c) = g[k];
    }
    if (d.p.dimension == SHCFD_DIM_AXISYMMETRIC && i >= 1 && i <= ni && j >= 1 && j <= nj) {
        // cell-centred metrics of the hoop stress (viscous/no-nV/q2d/sch_viscous.f90:10-28)
        double dxdxi = (XC(i + 1, j) - XC(i - 1, j)) * 0.5;
        double dydxi = (RC(i + 1, j) - RC(i - 1, j)) * 0.5;
        double dxdeta = (XC(i, j + 1) - XC(i, j - 1)) * 0.5;
        double dydeta = (RC(i, j + 1) - RC(i, j - 1)) * 0.5;
        double det = (dxdxi * dydeta) - (dydxi * dxdeta);
        FLD(d, SHCFD_F_GEOJACC, 0, c) = dydeta / det;
        FLD(d, SHCFD_F_GEOJACC, 2, c) = -dxdeta / det;
        FLD(d, SHCFD_F_GEOJACC, 1, c) = -dydxi / det;
        FLD(d, SHCFD_F_GEOJACC, 3, c) = dxdxi / det;
    }
}

}  // namespace shcfd
