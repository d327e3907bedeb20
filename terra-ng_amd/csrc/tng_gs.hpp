// Two-pass Gram-Schmidt of the FGMRES Arnoldi step (kernels_gs.hip).
#pragma once

#include <cstdint>

#include "tng_internal.hpp"

namespace tng {

// Basis of Stokes vectors: V (device memory) holds 4 component pointers per
// slot (3 velocity components, then the pressure); slots 0..j-1 are the final
// basis, slot j the tentative vector u_j, slot j+1 receives u_{j+1}.
struct GsArgs {
    double *const *V;
    const double *w[4];  // K z_j
    const uint8_t *mask_v, *mask_p;
    int64_t nv, np;  // node copies of the velocity and pressure meshes
    int j;
};

// q_j = (u_j - sum s_i q_i) inv_alpha,  u_{j+1} = (w - sum c_i q_i - r u_j) inv_nu
struct GsCoef {
    double s[32], c[32];
    double inv_alpha, inv_nu, r;
};

// out[0 .. 2B+2] (B = gs_bucket(j)): Q^T u (B slots, j used), Q^T w (B slots),
// u.u, u.w, w.w
int gs_bucket(int j);
size_t gs_partials_doubles();
void gs_dots(const GsArgs &a, double *partials, double *out, hipStream_t s);
void gs_update(const GsArgs &a, const GsCoef &cf, hipStream_t s);

// y += sum_{i < k} c_i X_i over Stokes vectors (X: device table of 4 component
// pointers per vector; y: device table of 4 component pointers)
struct GsAxpy {
    double c[8];
};
void gs_axpy(double *const *X, int k, const GsAxpy &cf, double *const *y, int64_t nv, int64_t np, hipStream_t s);
// v = e_axis x X, the rigid rotation about coordinate axis `axis`
void launch_rotation_field(tng_vector &v, int axis, hipStream_t s);

}  // namespace tng
