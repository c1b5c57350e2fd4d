// scf_driver.cpp -- rsg_scf_run: the SCF loop of calc_scf
// (scf_module.f90:183-481) in native code, over the library's own entry
// points and in the order the Python driver (scf.py) uses, so that a whole
// run makes no round trip to its caller.  Fermi-Dirac occupations
// (fermi_module.f90, mp_version = -1) and the entropy term
// (fermi_dirac_module.f90:325-350) are computed here on the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "context.hpp"
#include "solver.hpp"

namespace {

// Occupations 2 f((e - mu) / kT) with mu by bisection so that they sum to
// nel; returns false when the bands cannot hold nel electrons.
bool fermi_dirac(const std::vector<double> &esp, double nel, double kT, std::vector<double> &occ, double *mu_out) {
  const double w = 2.0;
  auto term = [&](double e, double mu) {
    const double x = std::min(700.0, std::max(-700.0, (e - mu) / kT));
    return w / (1.0 + std::exp(x));
  };
  auto count = [&](double mu) {
    double c = 0.0;
    for (double e : esp) c += term(e, mu);
    return c;
  };
  const double emin = *std::min_element(esp.begin(), esp.end()), emax = *std::max_element(esp.begin(), esp.end());
  double lo = emin - 50.0 * kT - 1.0, hi = emax + 50.0 * kT + 1.0;
  if (count(hi) < nel - 1e-12) return false;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count(mid) > nel)
      hi = mid;
    else
      lo = mid;
    if (hi - lo < 1e-15 * std::max(1.0, std::fabs(mid))) break;
  }
  const double mu = 0.5 * (lo + hi);
  occ.resize(esp.size());
  for (size_t i = 0; i < esp.size(); ++i) occ[i] = term(esp[i], mu);
  *mu_out = mu;
  return true;
}

// kT sum over bands of 2 [-f ln f - (1 - f) ln(1 - f)], f = occ / 2
double entropy_fd(const std::vector<double> &occ, double kT) {
  double s = 0.0;
  for (double o : occ) {
    const double f = 0.5 * o, g = 1.0 - f;
    if (f > 0.0) s -= f * std::log(f);
    if (g > 0.0) s -= g * std::log(g);
  }
  return 2.0 * kT * s;
}

}  // namespace

extern "C" int rsg_scf_run(rsg_ctx *ctx, const rsg_scf_params *p, rsg_scf_summary *out, double *etot_history,
                           double *esp_out) {
  if (!ctx || !p || !out) return RSG_ERR_ARG;
  if (p->nband < 1 || p->max_iter < 1 || p->ncg < 0 || !(p->ekbt > 0.0) || (p->imix != 10 && p->imix != 30))
    return ctx->fail(RSG_ERR_ARG, "scf_run: nband, max_iter >= 1, ncg >= 0, ekbt > 0 and imix 10/30 required");
  const auto t_start = std::chrono::steady_clock::now();
  int rc;
  double eewald = 0.0, nel = 0.0;
  if ((rc = rsg::scf_constants(ctx, &eewald, &nel))) return rc;
  const int natom = rsg::scf_atom_count(ctx);
  if (natom < 1) return ctx->fail(RSG_ERR_STATE, "scf_run: local potential not set");
  const int nb = p->nband;
  // set-up (scf.py SCF.__init__): density, potential, start bands, GS
  double eh_in = 0.0, exc_in = 0.0;
  if ((rc = rsg_init_density(ctx, p->init_width))) return rc;
  if ((rc = rsg_hartree_xc(ctx, RSG_FIELD_RHO, &eh_in, &exc_in))) return rc;
  if ((rc = rsg_update_vloc(ctx))) return rc;
  const bool learned = p->pc_kind == 5 || p->pc_kind == 6;
  if (learned && (rc = rsg_lpc_set_potential(ctx, nullptr))) return rc;
  if ((rc = rsg_alloc_bands(ctx, nb))) return rc;
  if ((rc = rsg_random_bands(ctx, 0, nb, p->seed))) return rc;
  if ((rc = rsg_gram_schmidt(ctx, 1, nb))) return rc;
  if ((rc = rsg_mix_init_imix(ctx, p->mmix, p->beta, p->imix))) return rc;
  if ((rc = rsg_set_precond_storage(ctx, p->precond_bits == 16 ? 16 : 32))) return rc;
  const bool store32 = p->mixed && p->fp32_store;
  std::vector<double> esp(static_cast<size_t>(nb)), res(static_cast<size_t>(nb)), occ;
  int prec = p->mixed ? 32 : 64, prev_prec = 0, stall = 0, fp32_its = 0;
  const double switch_below = p->fp64_below > 0.0 ? p->fp64_below : 2.0 * p->etol_per_atom;
  const double inf = std::numeric_limits<double>::infinity();
  double best = inf, e_prev = 0.0, etot = 0.0, fene = 0.0, efermi = 0.0;
  bool have_prev = false, conv = false;
  int it = 0;
  for (it = 1; it <= p->max_iter; ++it) {
    // FP32 iterations on the FP32 band store, FP64 ones on the FP64 store
    if (store32 && (rc = rsg_set_band_precision(ctx, prec))) return rc;
    // eigensolver half (scf_module.f90:231-268)
    if ((rc = rsg_subspace_diag(ctx, esp.data()))) return rc;
    rc = prec == 32 ? rsg_conjugate_gradient_f32(ctx, p->ncg, p->pc_kind, p->mloop, esp.data(), res.data())
                    : rsg_conjugate_gradient(ctx, p->ncg, p->pc_kind, p->mloop, esp.data(), res.data());
    if (rc) return rc;
    if ((rc = rsg_gram_schmidt(ctx, 1, nb))) return rc;
    if ((rc = rsg_esp_calc(ctx, esp.data()))) return rc;
    // occupations and energies (total_energy_module.f90:61-428)
    if (!fermi_dirac(esp, nel, p->ekbt, occ, &efermi))
      return ctx->fail(RSG_ERR_ARG, "scf_run: not enough bands for the electron count");
    double eeig = 0.0;
    for (int n = 0; n < nb; ++n) eeig += occ[size_t(n)] * esp[size_t(n)];
    double eloc_in = 0.0, eion_in = 0.0, eh = 0.0, exc = 0.0, eloc = 0.0, eion = 0.0;
    if ((rc = rsg_density_integrals(ctx, RSG_FIELD_RHO, &eloc_in, &eion_in))) return rc;
    if ((rc = rsg_calc_density(ctx, occ.data()))) return rc;
    if ((rc = rsg_hartree_xc(ctx, RSG_FIELD_RHO_OUT, &eh, &exc))) return rc;
    if ((rc = rsg_density_integrals(ctx, RSG_FIELD_RHO_OUT, &eloc, &eion))) return rc;
    etot = eeig - eloc + eh + exc + eion + eewald;
    fene = etot - entropy_fd(occ, p->ekbt);
    const double de = have_prev ? std::fabs(etot - e_prev) / natom : inf;
    if (etot_history) etot_history[it - 1] = etot;
    const int this_prec = prec;
    if (prec == 32) {
      ++fp32_its;
      if (de < best) {
        best = de;
        stall = 0;
      } else {
        ++stall;
      }
      if (de < switch_below || stall >= p->fp32_stall) prec = 64;  // FP32 never ends the loop
    }
    conv = this_prec == 64 && prev_prec == 64 && de < p->etol_per_atom;
    prev_prec = this_prec;
    if (!conv && it < p->max_iter) {
      // next input density and potential (scf_module.f90:446-456)
      if ((rc = rsg_mix_pulay(ctx))) return rc;
      if ((rc = rsg_normalize_density(ctx))) return rc;
      if ((rc = rsg_hartree_xc(ctx, RSG_FIELD_RHO, &eh_in, &exc_in))) return rc;
      if ((rc = rsg_update_vloc(ctx))) return rc;
      if (learned && (rc = rsg_lpc_set_potential(ctx, nullptr))) return rc;
    }
    e_prev = etot;
    have_prev = true;
    if (conv) break;
  }
  if (store32 && (rc = rsg_set_band_precision(ctx, 64))) return rc;
  out->converged = conv ? 1 : 0;
  out->iterations = std::min(it, p->max_iter);
  out->fp32_iterations = fp32_its;
  out->etot = etot;
  out->fene = fene;
  out->efermi = efermi;
  out->eewald = eewald;
  out->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (esp_out) std::copy(esp.begin(), esp.end(), esp_out);
  return RSG_OK;
}
