// Species mass-flux chain on the device (lmm_species.hip): the per-cell
// algebra of mass_flux_utilities.f90 / compute_mixture_properties.f90 /
// matrix_utilities.f90 and the face fluxes of diffusive_mass_fluxdiv.f90 and
// stochastic_mass_fluxdiv.f90.  Fields are component-major deduplicated
// periodic arrays (component q of cell t at q*N + t); species matrices are
// column-major (ns, ns), entry (r, c) in component r + ns*c.
#pragma once

#include <hip/hip_runtime.h>

#include "lmm_common.h"

namespace lmm {

constexpr int SP_MAX = 8;

// Knobs of probin_multispecies / probin_common (passed by value to kernels).
struct SpParams {
  int ns;
  double molmass[8], rhobar[8];
  double Dbar[28], Dtherm[8], Hoff[28], Hdiag[8];
  double k_B, vcm, ftol;
  int ideal, noniso, lapack, inverse_type, chi_iter, correct, baro;
};

// Outputs of the per-cell chain kernel that feeds the fused face kernel.
struct SpCellOut {
  double *rho;    // corrected for the cell algebra, written back as rho_c - drho
  double *xm;     // molarconc [ns]
  double *rw;     // rhoWchi [ns*ns]
  double *gama;   // Gama [ns*ns], nullptr for an ideal mixture (identity)
  double *zeta;   // zeta_by_Temp [ns], nullptr when isothermal
  double *lpk;    // Cholesky factor of L-Onsager, lower triangle by columns [ns(ns+1)/2], or nullptr
  double *baro;   // barodiffusion coefficient [ns], or nullptr
};

// Inputs of the fused face kernel.
struct SpFaceIn {
  const double *xm, *rw, *gama, *zeta, *temp, *baro, *lpk;
  const double *gpb[3];     // gradp_baro faces (barodiffusion) or nullptr
  const double *W;          // noise stages [n_rngs][dim][ns][N] or nullptr
  const double *weights;    // device [n_rngs]
  int n_rngs;
  double variance;
};

// ---- reference routines one by one (each returns 0 or an lmm status) ----
int sp_correct_rho(hipStream_t s, const Grid &g, const SpParams &sp, double *rho, double *drho);
int sp_cons_to_prim(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, double *rhotot,
                    double *molarconc, double *molmtot);
int sp_mixture(hipStream_t s, const Grid &g, const SpParams &sp, double *D_bar, double *D_therm, double *H);
int sp_gama(hipStream_t s, const Grid &g, const SpParams &sp, const double *molarconc, const double *H,
            double *Gama);
int sp_chi(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rhotot,
           const double *molarconc, const double *D_bar, const double *D_therm, const double *Temp, double *chi,
           double *zeta);
int sp_rhowchi(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *chi,
               double *rhoWchi);
int sp_lonsager(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rhotot,
                const double *molmtot, const double *chi, double *L);
// diffusive_mass_flux (with the temperature piece when zeta and Temp are
// given, the barodiffusion piece when gpb is given, its cell coefficient
// computed into baro[ns*N]): face fluxes into flux[d*ns*N + s*N + t],
// flux_total[d] incremented, then the divergence.  Gama null: identity.
int sp_diffusive(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *molarconc,
                 const double *rhoWchi, const double *Gama, const double *Temp, const double *zeta,
                 const double *const gpb[3], double *baro, double *flux, double *const flux_total[3], double *div);
// stochastic_mass_fluxdiv from full (ns, ns) Cholesky factors per cell
int sp_stochastic(hipStream_t s, const Grid &g, const SpParams &sp, const double *L, const double *W, int n_rngs,
                  const double *weights, double variance, double *flux, double *const flux_total[3], double *div,
                  int increment);
// compute_mass_fluxdiv: one cell pass, then one face pass for both fluxes
// (the diffusive one alone when W is null) and one divergence pass; flux is
// scratch of 2*dim*ns*N values (dim*ns*N without W); flux_total is
// overwritten with (0 + F_diff) + F_stoch, the reference's setval then adds
int sp_chain_cell(hipStream_t s, const Grid &g, const SpParams &sp, const double *Temp, const SpCellOut &o);
int sp_chain_face(hipStream_t s, const Grid &g, const SpParams &sp, const SpFaceIn &in, double *flux,
                  double *diff_div, double *stoch_div, double *const flux_total[3]);

}  // namespace lmm
