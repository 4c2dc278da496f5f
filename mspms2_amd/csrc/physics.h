// physics.h -- pair terms shared by the gfx950 kernels.
//
// FP64 versions keep the reference's expression order with contraction
// disabled, so that they round like the reference's x86-64 build; the libm
// calls become ROCm device-library calls (erfc, exp: ~1 ulp).
// FP32 versions are the mixed-precision pair arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#define MSPMS_PI 3.141592653589793  // src/defs.h:100

// ---- FP64 ---------------------------------------------------------------
// ljfrc (src/ljfrc.c:21-71)
__host__ __device__ inline void lj64(double rijsq, double sigmaij, double epsilonij, int switch_on,
                            double rcutoffsq, double rcutonsq, double roff2_ron2_cube,
                            double shift1, double* uij, double* fij, double* ushift) {
#pragma clang fp contract(off)
    double sw = 0.0;
    if (switch_on) {
        if (rijsq < rcutonsq)
            sw = 1.0;
        else
            sw = (rcutoffsq - rijsq) * (rcutoffsq - rijsq) * (rcutoffsq + 2.0 * rijsq - 3.0 * rcutonsq) /
                 roff2_ron2_cube;
    }
    double r_rijsq = 1.0 / rijsq;
    double s2 = sigmaij * sigmaij;
    double r6 = s2 * r_rijsq;
    r6 = r6 * r6 * r6;
    double r12 = r6 * r6;
    double u = 4.0 * epsilonij * (r12 - r6);
    if (switch_on) {
        if (rijsq < rcutonsq) {
            *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq;
        } else {
            *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq * sw -
                   u * 12.0 * (rcutoffsq - rijsq) * (rcutonsq - rijsq) / roff2_ron2_cube;
            u *= sw;
        }
        *ushift = 0.0;
    } else {
        *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq;
        double s6 = s2 * s2 * s2;
        *ushift = epsilonij * s6 * (s6 * shift1 - 1.0);
    }
    *uij = u;
}

// Parameter derivatives of the ljfrc energy: du/dsigma_ij and du/deps_ij of
// u = 4 eps (s^12 - s^6) [* sw(r)], s = sigma / r; the switch factor does
// not depend on the parameters.
__host__ __device__ inline void lj64_dparam(double rijsq, double sigmaij, double epsilonij, int switch_on,
                                   double rcutoffsq, double rcutonsq, double roff2_ron2_cube,
                                   double* du_dsigma, double* du_deps) {
    double sw = 1.0;
    if (switch_on && !(rijsq < rcutonsq))
        sw = (rcutoffsq - rijsq) * (rcutoffsq - rijsq) * (rcutoffsq + 2.0 * rijsq - 3.0 * rcutonsq) /
             roff2_ron2_cube;
    double r6 = sigmaij * sigmaij / rijsq;
    r6 = r6 * r6 * r6;
    const double r12 = r6 * r6;
    *du_deps = 4.0 * (r12 - r6) * sw;
    *du_dsigma = 24.0 * epsilonij * (2.0 * r12 - r6) / sigmaij * sw;
}

// d eps_ij / d eps_i of the mixing rule eps_ij = sqrt(eps_i eps_j): infinite
// at eps_i = 0 < eps_j, where the square root has no derivative
__host__ __device__ inline double deps_mix(double eij, double eps_j) {
    if (eij > 0.0) return 0.5 * eps_j / eij;
    return eps_j == 0.0 ? 0.0 : __builtin_huge_val();
}

// ewald_real_frc (src/elecfrc.c:17-33)
__host__ __device__ inline void ewald_real64(double rijsq, double qq, double C, double kappa, double* uij,
                                    double* fij) {
#pragma clang fp contract(off)
    double rij = sqrt(rijsq);
    double ut = C * qq / rij;
    *uij = ut * erfc(kappa * rij);
    double t1 = kappa * rij;
    double t2 = ut * erfc(t1);
    *fij = (t2 + ut * 2.0 * t1 * exp(-t1 * t1) / sqrt(MSPMS_PI)) / rijsq;
}

// ewald_excl_frc (src/elecfrc.c:35-46)
__host__ __device__ inline void ewald_excl64(double rijsq, double qq, double C, double* uij, double* fij) {
#pragma clang fp contract(off)
    double rij = sqrt(rijsq);
    *uij = C * qq / rij;
    *fij = -*uij / rijsq;
}

// xatom_coulomb_frc (src/elecfrc.c:159-170): bare Coulomb, u = C qq / r,
// F_i = (u / r^2) r_ij
__host__ __device__ inline void coulomb64(double rijsq, double qq, double C, double* uij,
                                          double* fij) {
#pragma clang fp contract(off)
    double rij = sqrt(rijsq);
    *uij = C * qq / rij;
    *fij = *uij / rijsq;
}

// wolf_real_frc (src/elecfrc.c:48-60)
__host__ __device__ inline void wolf_real64(double rijsq, double qq, double C, double kappa, double rc,
                                   double v1, double v2, double f1, double f2, double* uij,
                                   double* fij) {
#pragma clang fp contract(off)
    double rij = sqrt(rijsq);
    *uij = C * qq * (erfc(kappa * rij) / rij + v1 + v2 * (rij - rc));
    *fij = C * (qq / rij) * (erfc(kappa * rij) / rijsq + f1 * exp(-(kappa * rij) * (kappa * rij)) / rij +
                             f2);
}

// ---- FP32 ---------------------------------------------------------------
// erfc(x) = exp(-x^2) * P(t), t = 1/(1 + x/2), x >= 0.  P is a degree-9
// weighted least-squares fit of the scaled complementary error function
// exp(x^2) erfc(x) on [0, 12] (max relative error 4e-8 in exact
// arithmetic, 3.3e-7 evaluated in FP32).  exp(-x^2) is shared with the
// Ewald / Wolf force term.
__device__ inline float erfc_scaled_poly(float x) {
    // hardware reciprocal (v_rcp_f32, ~1 ulp): the fit's error stays ~4e-7
    float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
    float p = -5.734027581e-02f;
    p = fmaf(p, t, 2.758059454e-01f);
    p = fmaf(p, t, -4.681025672e-01f);
    p = fmaf(p, t, 2.547098534e-01f);
    p = fmaf(p, t, -9.007769349e-03f);
    p = fmaf(p, t, 1.942188194e-01f);
    p = fmaf(p, t, 2.456883800e-01f);
    p = fmaf(p, t, 2.818937015e-01f);
    p = fmaf(p, t, 2.821360283e-01f);
    p = fmaf(p, t, -2.151000532e-06f);
    return p;
}

#define MSPMS_TWO_OVER_SQRTPI_F 1.1283791670955126f
