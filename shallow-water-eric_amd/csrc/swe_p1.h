// swe_p1.h -- P1 reference element and per-element geometry of the set-up, shared by the host
// assembly (swe_setup.cpp) and the device assembly (swe_setup_device.hip): one definition, so the
// two produce the same bits.
//
//  * geometry of gauss_points_2d_p1 (SFEMaNS/GAUSS_POINTS_OBJET/gauss_points_2d_p1.f90:106-132)
//    with the 3-point rule of element_2d (347-393).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace swe_p1 {

struct P1Rule {
    double ww[3][3];   // ww[n][l]
    double dd[2][3];   // reference gradients dd[k][n]
    double pp;         // weight 1/6
    __host__ __device__ P1Rule() {
        const double one = 1, two = 2, three = 3, six = 6;
        const double xx[3] = {one / six, two / three, one / six};
        const double yy[3] = {one / six, one / six, two / three};
        for (int l = 0; l < 3; ++l) {
            ww[0][l] = one - xx[l] - yy[l];
            ww[1][l] = xx[l];
            ww[2][l] = yy[l];
        }
        dd[0][0] = -one; dd[1][0] = -one;
        dd[0][1] = one;  dd[1][1] = 0.0;
        dd[0][2] = 0.0;  dd[1][2] = one;
        pp = one / six;
    }
};

// dw[k][n] and rj[l] of one element (the P1 gradient is the same at its three Gauss points)
struct ElemGeom {
    double dw[2][3][3];
    double rj[3];
};

__host__ __device__ inline void element_geometry(const P1Rule& R, const double* rr, const int* nd, ElemGeom& G) {
    for (int l = 0; l < 3; ++l) {
        double dr[2][2];
        for (int k = 0; k < 2; ++k)
            for (int k1 = 0; k1 < 2; ++k1) {
                double s = 0.0;
                for (int n = 0; n < 3; ++n) s = s + rr[2 * (size_t)nd[n] + k] * R.dd[k1][n];
                dr[k][k1] = s;
            }
        double rjac = dr[0][0] * dr[1][1] - dr[0][1] * dr[1][0];
        for (int n = 0; n < 3; ++n) {
            G.dw[0][n][l] = (+R.dd[0][n] * dr[1][1] - R.dd[1][n] * dr[1][0]) / rjac;
            G.dw[1][n][l] = (-(R.dd[0][n] * dr[0][1]) + R.dd[1][n] * dr[0][0]) / rjac;
        }
        G.rj[l] = rjac * R.pp;
    }
}

}  // namespace swe_p1
