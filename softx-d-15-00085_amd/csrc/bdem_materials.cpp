/*
 * bdem_materials.cpp - setup-time spring constants (host C++, no device).
 *
 * Restates Set_SpringConstants{Particle,Surface,DSurface}_{KN,TC}
 * (Input/KSimulationData.cpp:887-1071) and the loops that call them
 * (:276-292 for pairs, :1102-1168 per type): double arithmetic, results
 * stored in the POBJ's float fields, in the reference's call order (pairs
 * row by row, so the time-of-contact mode's in-place friction averaging is
 * order-dependent exactly as written).
 */
#include <cmath>
#include "bdem_abi.h"

namespace {

constexpr int PP_SLOTS = 6;   /* POBJ::PP[6], KSimulationObjects.h:290 */

double shifted(double cor) {
  /* the fitted shift for 0.3 <= e <= 0.9 (:893-896, :968-971) */
  if (cor >= 0.30 && cor <= 0.90) cor -= 0.038080;
  return cor;
}

/* Set_SpringConstantsParticleKN (:955-1017) */
void particle_kn(bdem_POBJ* P, int i, int j, double cor_i, double cor_j) {
  double COR = cor_i;
  double m = 1.0 / ((1.0 / P[i].mass) + (1.0 / P[i].mass));   /* particle i's mass twice, as written */
  const double Pi = 2.0 * std::asin(1.0);
  if (i == j) {
    if (COR < 1.000) {
      COR = shifted(COR);
      P[i].PP[j].Cn = -2.0 * std::log(COR) * std::sqrt(P[i].PP[j].Kn * m) /
                      (std::sqrt((std::log(COR) * std::log(COR) + Pi * Pi)));
    } else {
      P[i].PP[j].Cn = 0.0f;
    }
  }
  if (i != j) {
    m = 1.0 / ((1.0 / P[i].mass) + (1.0 / P[j].mass));
    COR = 0.50f * (cor_i + cor_j);
    COR = shifted(COR);
    P[i].PP[j].Kn = 0.50f * (P[i].PP[i].Kn + P[j].PP[j].Kn);
    P[i].PP[j].Fric_static = 0.50 * (P[i].PP[i].Fric_static + P[j].PP[j].Fric_static);
    P[i].PP[j].Fric_kinetic = 0.50 * (P[i].PP[i].Fric_kinetic + P[j].PP[j].Fric_kinetic);
    if (COR < 1.000)
      P[i].PP[j].Cn = -2.0 * std::log(COR) * std::sqrt(P[i].PP[j].Kn * m) /
                      (std::sqrt((std::log(COR) * std::log(COR) + Pi * Pi)));
    else
      P[i].PP[j].Cn = 0.0f;
  }
}

/* Set_SpringConstantsParticle_TC (:887-918) */
void particle_tc(bdem_POBJ* P, int i, int j, double cor_i, double cor_j, double tc) {
  double COR = 0.50f * (cor_i + cor_j);
  double m = 1.0 / ((1.0 / P[i].mass) + (1.0 / P[j].mass));
  COR = shifted(COR);
  if (i == j) m = P[i].mass;   /* the particle's own mass, not the reduced mass */
  const double Pi = 2.0 * std::asin(1.0);
  P[i].PP[j].Cn = 2.0 * m * (-std::log(COR) / (tc));
  P[i].PP[j].Kn = (m / (tc * tc)) * (std::log(COR) * std::log(COR) + Pi * Pi);
  /* PP[0] holds the value read from the deck (and may already be averaged) */
  P[i].PP[j].Fric_static = 0.50 * (P[i].PP[0].Fric_static + P[j].PP[0].Fric_static);
  P[i].PP[j].Fric_kinetic = 0.50 * (P[i].PP[0].Fric_kinetic + P[j].PP[0].Fric_kinetic);
}

/* Set_SpringConstantsSurface_KN / DSurfaceKN (:1019-1071) */
float wall_cn_kn(double kn, double m, double COR) {
  const double Pi = 2.0 * std::asin(1.0);
  if (COR < 1.000) {
    COR = shifted(COR);
    return -2.0 * std::log(COR) * std::sqrt(kn * m) / (std::sqrt((std::log(COR) * std::log(COR) + Pi * Pi)));
  }
  return 0.0f;
}

}  // namespace

namespace bdem {

/* The deck reader's order: per type as each POBJ is read (:1102-1168, with
 * that type's own surface and lifter contact times), then the pair matrix
 * row by row (:276-292, with one pair contact time: the reader keeps only
 * the last type's). */
int derive_springs(bdem_POBJ* pobj, int num_pobj, int mode, const double* cor_pp, const double* cor_surface,
                   const double* cor_dsurface, double tc_pp, const double* tc_ps, const double* tc_pd) {
  if (!pobj || !cor_pp || !cor_surface || !cor_dsurface) return BDEM_EINVAL;
  if (num_pobj < 1 || num_pobj > PP_SLOTS) return BDEM_EINVAL;
  if (mode != BDEM_SPRING_TC && mode != BDEM_SPRING_KN) return BDEM_EINVAL;
  for (int i = 0; i < num_pobj; ++i)
    if (!(pobj[i].mass > 0.0f)) return BDEM_EINVAL;
  if (mode == BDEM_SPRING_TC) {
    if (!tc_ps || !tc_pd || !(tc_pp > 0.0)) return BDEM_EINVAL;
    for (int i = 0; i < num_pobj; ++i)
      if (!(tc_ps[i] > 0.0 && tc_pd[i] > 0.0)) return BDEM_EINVAL;
  }
  for (int i = 0; i < num_pobj; ++i) {
    bdem_POBJ& p = pobj[i];
    const double m = p.mass;
    if (mode == BDEM_SPRING_TC) {
      const double Pi = 2.0 * std::asin(1.0);
      double cs = shifted(cor_surface[i]);
      p.surface_Cn = 2.0 * m * (-std::log(cs) / (tc_ps[i]));
      p.surface_Kn = (m / (tc_ps[i] * tc_ps[i])) * (std::log(cs) * std::log(cs) + Pi * Pi);
      double cd = shifted(cor_dsurface[i]);
      p.Dsurface_Cn = 2.0 * m * (-std::log(cd) / (tc_pd[i]));
      p.Dsurface_Kn = (m / (tc_pd[i] * tc_pd[i])) * (std::log(cd) * std::log(cd) + Pi * Pi);
    } else {
      p.surface_Cn = wall_cn_kn(p.surface_Kn, m, cor_surface[i]);
      p.Dsurface_Cn = wall_cn_kn(p.Dsurface_Kn, m, cor_dsurface[i]);
    }
  }
  for (int i = 0; i < num_pobj; ++i)
    for (int j = 0; j < num_pobj; ++j) {
      if (mode == BDEM_SPRING_TC) particle_tc(pobj, i, j, cor_pp[i], cor_pp[j], tc_pp);
      else particle_kn(pobj, i, j, cor_pp[i], cor_pp[j]);
    }
  return BDEM_OK;
}

}  // namespace bdem

extern "C" int bdem_derive_spring_constants(bdem_POBJ* pobj, int num_pobj, int mode, const double* cor_pp,
                                            const double* cor_surface, const double* cor_dsurface,
                                            double tc_pp, double tc_ps, double tc_pd) {
  double ps[PP_SLOTS], pd[PP_SLOTS];
  for (int i = 0; i < PP_SLOTS; ++i) { ps[i] = tc_ps; pd[i] = tc_pd; }
  return bdem::derive_springs(pobj, num_pobj, mode, cor_pp, cor_surface, cor_dsurface, tc_pp, ps, pd);
}
