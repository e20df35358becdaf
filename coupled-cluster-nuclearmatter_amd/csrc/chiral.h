// Host side of the built-in Delta-full chiral NN: the constants of one chiral
// order and cutoff, and the tables the device evaluator reads (interaction.h).
#pragma once
#include <vector>

#include "interaction.h"

namespace ccgpu {

struct ChiralModel {
  int order = 3;          // 0 LO, 2 NLO, 3 NNLO
  double cutoff = 450.0;  // regulator cutoff, MeV
  double nexp = 3.0;      // regulator exponent
  double c1 = 0, c2 = 0, c3 = 0, c4 = 0;   // pi-N LECs, MeV^-1
  double cs[3] = {0, 0, 0}, ct[3] = {0, 0, 0};   // LO contacts per pair charge (pp, np, nn)
  double cnlo[7] = {0, 0, 0, 0, 0, 0, 0};        // NLO contacts C1..C7
  double cg0[2] = {0, 0}, cg1[2] = {0, 0};       // isospin CG of an np pair, T = 0 and 1
};

// Constants of init_chp_constants (chiral_module_andreas_with_delta.f90:122-808)
// for order in {0, 2, 3} and cutoff in {394, 450, 500} MeV.  ci (4 values in
// GeV^-1, or null for the defaults of the order) replaces the NN c1..c4.
ChiralModel chiral_model(int order, double cutoff, const double *ci);

// Loop functions of the momentum transfer for m = 0..mmax with
// q^2 = (qunit^2) m, kChiralRad values per m (layout in interaction.h).
std::vector<double> chiral_radial_table(const ChiralModel &c, double qunit, int mmax);

// Regulator x relativity half-factor per pair charge (pp, np, nn) and
// j = 0..jmax with p^2 = (qunit^2) j: 3 x (jmax + 1) values.
std::vector<double> chiral_side_table(const ChiralModel &c, double qunit, int jmax);

// Constants of the NNLO three-nucleon force for the chiral model c: cE, cD
// (zero below NNLO, as ccm_main.f90:33-36), its c1, c3, c4 (those of the NN
// at NNLO, zero below; init_chp_constants) shifted by the Delta, and the
// regulator cutoff (MeV) and power of freg_3nfnonlocal.  Momenta pointers and
// the volume scale are set by the caller.
Chiral3Dev chiral3_constants(const ChiralModel &c, double cE, double cD, double cutoff3, int nexp3);

}  // namespace ccgpu
