/* CPU restatement of CCAKE's pre-inverted EoS for the flat (s, nB) format
 * (EoS_Interpolator, include/eos_interpolator.h, src/eos_interpolator.cpp),
 * used when online_inverter_enabled is false:
 *   - load_flat_table (291-527): unit conversion, log-uniform axes, dw/ds and
 *     dw/dB from the baryon-only Schur complement of the susceptibilities;
 *   - fill_thermodynamics (531-627) through interpolate4D's bilinear branch
 *     (639-739): the per-step lookup from (s, rhoB);
 *   - query_thermo_host (751-811) and invert_e_to_s_host (815-861): the
 *     initial e -> s inversion;
 *   - efreeze_at_T (865-892).
 * Test infrastructure only (the parity checker), like the rest of oracle/.
 */
#include "orc_internal.h"

#define PV(p, k, is, iB) ((p)->v[k][(size_t)(is)*(p)->NB + (iB)])

static void preinv_free_arrays(orc_preinv* p)
{
  free(p->s_axis); free(p->nB_axis);
  for (int k = 0; k < PV_NUM; ++k) free(p->v[k]);
}

void orc_preinv_destroy(orc_preinv* p)
{
  if (!p) return;
  preinv_free_arrays(p);
  free(p);
}

/* log_step of load_flat_table: log10 start and stride, or a rejection when
 * the axis is not uniform in log10 to 1e-6. */
static int log_step(const double* a, int n, double* l0_out, double* step_out)
{
  const double l0 = log10(a[0]);
  const double l1 = log10(a[n-1]);
  const double step = (l1 - l0)/(n - 1);
  double max_dev = 0.0;
  for (int i = 1; i + 1 < n; ++i) {
    const double d = log10(a[i+1]) - log10(a[i]) - step;
    if (fabs(d) > max_dev) max_dev = fabs(d);
  }
  if (max_dev > 1e-6) return 0;
  *l0_out = l0; *step_out = step;
  return 1;
}

orc_preinv* orc_preinv_create(const orc_eos_desc* d)
{
  const int Ns = d->dims[0], NB = d->dims[1];
  if (Ns < 2 || NB < 2 || !d->table) return NULL;
  const size_t nt = (size_t)Ns*NB;
  const double* in[12];
  for (int k = 0; k < 12; ++k) in[k] = d->table + (size_t)k*nt;
  /* s, nB, T, muB, muS, muQ, e, p, cs2, chiTT, chiTB, chiBB */
  orc_preinv* p = (orc_preinv*)calloc(1, sizeof(orc_preinv));
  p->Ns = Ns; p->NB = NB;
  p->s_axis = (double*)malloc(sizeof(double)*Ns);
  p->nB_axis = (double*)malloc(sizeof(double)*NB);
  for (int k = 0; k < PV_NUM; ++k) p->v[k] = (double*)malloc(sizeof(double)*nt);
  const double inv_hc1 = 1.0/ORC_HBARC_MEVFM;
  const double inv_hc2 = inv_hc1*inv_hc1;
  const double inv_hc3 = inv_hc2*inv_hc1;
  const double inv_hc4 = inv_hc2*inv_hc2;
  for (int is = 0; is < Ns; ++is) p->s_axis[is] = in[0][(size_t)is*NB + 0]*inv_hc3;
  for (int iB = 0; iB < NB; ++iB) p->nB_axis[iB] = in[1][iB]*inv_hc3;
  for (int is = 0; is < Ns; ++is) if (!(p->s_axis[is] > 0.0)) { orc_preinv_destroy(p); return NULL; }
  for (int iB = 0; iB < NB; ++iB) if (!(p->nB_axis[iB] > 0.0)) { orc_preinv_destroy(p); return NULL; }
  if (!log_step(p->s_axis, Ns, &p->ls0, &p->dls) || !log_step(p->nB_axis, NB, &p->lb0, &p->dlb)) {
    orc_preinv_destroy(p); return NULL;
  }
  const double EPS = 1e-30, TINY = 1e-25;
  for (int is = 0; is < Ns; ++is)
  for (int iB = 0; iB < NB; ++iB) {
    const size_t i = (size_t)is*NB + iB;
    const double T_c = in[2][i]*inv_hc1, muB_c = in[3][i]*inv_hc1;
    const double chiTT = in[9][i]*inv_hc2, chiTB = in[10][i]*inv_hc2, chiBB = in[11][i]*inv_hc2;
    const double s_c = p->s_axis[is], nB_c = p->nB_axis[iB];
    const double dentr_dt   = chiTT - (chiTB*chiTB)/(chiBB + EPS);
    const double dentr_dmub = chiTB - (chiTT*chiBB)/(chiTB + EPS);
    const double db_dt      = chiTB - (chiBB*chiTT)/(chiTB + EPS);
    const double db_dmub    = chiBB - (chiTB*chiTB)/(chiTT + EPS);
    double dwds_c = T_c + s_c/dentr_dt;
    if (fabs(nB_c) > TINY) dwds_c += nB_c/dentr_dmub;
    double dwdB_c = muB_c;
    if (fabs(nB_c) > TINY) dwdB_c += s_c/db_dt + nB_c/db_dmub;
    p->v[PV_T][i] = T_c;
    p->v[PV_MUB][i] = muB_c;
    p->v[PV_MUS][i] = in[4][i]*inv_hc1;
    p->v[PV_MUQ][i] = in[5][i]*inv_hc1;
    p->v[PV_E][i] = in[6][i]*inv_hc4;
    p->v[PV_P][i] = in[7][i]*inv_hc4;
    p->v[PV_CS2][i] = in[8][i];
    p->v[PV_DWDS][i] = dwds_c;
    p->v[PV_DWDB][i] = dwdB_c;
    p->v[PV_DWDQ][i] = 0.0;
    p->v[PV_DWDS_S][i] = 0.0;
  }
  return p;
}

/* fill_thermodynamics for the flat slot: S and Q axes degenerate, so
 * interpolate4D takes its bilinear (s, rhoB) branch in log10 coordinates.
 * out: T, muB, muS, muQ, e, p, cs2, dwds, dwdB, dwdS, dwdQ, w.
 * The reference clamps the cell index to N-1 and then reads node N at the
 * upper edge; here the cell is clamped to N-2, which gives the same value
 * at the last node and extrapolates the last cell beyond it. */
static void preinv_fill(const orc_preinv* p, double s_in, double rhoB_in, double out[12])
{
  const double s = fmax(.001, s_in);
  const double rhoB = fmax(.001, fabs(rhoB_in));
  const double pos0 = log10(fmax(s, 1e-30));
  const double pos1 = log10(fmax(rhoB, 1e-30));
  int i0 = (fabs(p->dls) < 1e-30) ? 0 : (int)floor((pos0 - p->ls0)/p->dls);
  int i1 = (fabs(p->dlb) < 1e-30) ? 0 : (int)floor((pos1 - p->lb0)/p->dlb);
  i0 = i0 < 0 ? 0 : (i0 > p->Ns - 2 ? p->Ns - 2 : i0);
  i1 = i1 < 0 ? 0 : (i1 > p->NB - 2 ? p->NB - 2 : i1);
  const double x0 = (pos0 - p->ls0 - i0*p->dls)/p->dls;
  const double x1 = (pos1 - p->lb0 - i1*p->dlb)/p->dlb;
  double f[PV_NUM];
  for (int k = 0; k < PV_NUM; ++k) {
    double acc = 0.0;
    for (int c = 0; c < 4; ++c) {
      double phi = 1.0;
      phi *= (c & 1) ? x0 : (1.0 - x0);
      phi *= (c & 2) ? x1 : (1.0 - x1);
      acc += phi*PV(p, k, i0 + (c & 1), i1 + ((c >> 1) & 1));
    }
    f[k] = acc;
  }
  out[0] = f[PV_T]; out[1] = f[PV_MUB]; out[2] = f[PV_MUS]; out[3] = f[PV_MUQ];
  out[4] = f[PV_E]; out[5] = f[PV_P]; out[6] = f[PV_CS2];
  out[7] = f[PV_DWDS]; out[8] = f[PV_DWDB]; out[9] = f[PV_DWDS_S]; out[10] = f[PV_DWDQ];
  out[11] = out[4] + out[5];
}

/* query_thermo_host: bilinear in (log10 s, log10 |rhoB|), both clamped into
 * the table.  out: T, muB, muS, muQ, e, p, cs2, dwds, dwdB, w. */
static void preinv_query(const orc_preinv* p, double s, double rhoB, double out[10])
{
  const int Ns = p->Ns, NB = p->NB;
  const double abs_rhoB = fmax(fabs(rhoB), p->nB_axis[0]);
  const double s_clamped = fmax(s, p->s_axis[0]);
  const double L_s_max = p->ls0 + p->dls*(Ns - 1);
  const double L_nB_max = p->lb0 + p->dlb*(NB - 1);
  const double L_s = fmin(fmax(log10(s_clamped), p->ls0), L_s_max);
  const double L_nB = fmin(fmax(log10(abs_rhoB), p->lb0), L_nB_max);
  const double xs = (L_s - p->ls0)/p->dls;
  const double xb = (L_nB - p->lb0)/p->dlb;
  int is = (int)floor(xs), iB = (int)floor(xb);
  is = is < 0 ? 0 : (is > Ns - 2 ? Ns - 2 : is);
  iB = iB < 0 ? 0 : (iB > NB - 2 ? NB - 2 : iB);
  const double a = xs - is, b = xb - iB;
#define BIL(k) ((1.0 - a)*(1.0 - b)*PV(p, k, is, iB) + a*(1.0 - b)*PV(p, k, is + 1, iB) \
                + (1.0 - a)*b*PV(p, k, is, iB + 1) + a*b*PV(p, k, is + 1, iB + 1))
  out[0] = BIL(PV_T); out[1] = BIL(PV_MUB); out[2] = BIL(PV_MUS); out[3] = BIL(PV_MUQ);
  out[4] = BIL(PV_E); out[5] = BIL(PV_P); out[6] = BIL(PV_CS2);
  out[9] = out[4] + out[5];
  out[7] = BIL(PV_DWDS); out[8] = BIL(PV_DWDB);
#undef BIL
}

/* invert_e_to_s_host: bisection on the s axis of e at fixed |rhoB|, then
 * log10-linear in s inside the bracketing cell.  Returns s. */
static double preinv_e_to_s(const orc_preinv* p, double e_input, double rhoB)
{
  const int Ns = p->Ns, NB = p->NB;
  const double L_nB_max = p->lb0 + p->dlb*(NB - 1);
  const double abs_rhoB = fmax(fabs(rhoB), p->nB_axis[0]);
  const double L_nB = fmin(log10(abs_rhoB), L_nB_max);
  const double xi_B = (L_nB - p->lb0)/p->dlb;
  int iB = (int)floor(xi_B);
  iB = iB < 0 ? 0 : (iB > NB - 2 ? NB - 2 : iB);
  const double b = xi_B - iB;
#define E_AT(is_) ((1.0 - b)*PV(p, PV_E, (is_), iB) + b*PV(p, PV_E, (is_), iB + 1))
  if (e_input <= E_AT(0)) return p->s_axis[0];
  if (e_input >= E_AT(Ns - 1)) return p->s_axis[Ns - 1];
  int lo = 0, hi = Ns - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi)/2;
    if (E_AT(mid) <= e_input) lo = mid; else hi = mid;
  }
  const double e_lo = E_AT(lo), e_hi = E_AT(hi);
#undef E_AT
  const double alpha = (e_hi > e_lo) ? (e_input - e_lo)/(e_hi - e_lo) : 0.0;
  const double L_lo = log10(p->s_axis[lo]), L_hi = log10(p->s_axis[hi]);
  return pow(10.0, L_lo + alpha*(L_hi - L_lo));
}

double orc_preinv_efreeze(const orc_preinv* p, double T_fm)
{
  const int Ns = p->Ns;
  const double T_lo_edge = PV(p, PV_T, 0, 0), T_hi_edge = PV(p, PV_T, Ns - 1, 0);
  if (T_fm <= T_lo_edge) return PV(p, PV_E, 0, 0);
  if (T_fm >= T_hi_edge) return PV(p, PV_E, Ns - 1, 0);
  int lo = 0, hi = Ns - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi)/2;
    if (PV(p, PV_T, mid, 0) <= T_fm) lo = mid; else hi = mid;
  }
  const double T_l = PV(p, PV_T, lo, 0), T_h = PV(p, PV_T, hi, 0);
  const double alpha = (T_h > T_l) ? (T_fm - T_l)/(T_h - T_l) : 0.0;
  return (1.0 - alpha)*PV(p, PV_E, lo, 0) + alpha*PV(p, PV_E, hi, 0);
}

/* mode 0: SPHWorkstation::update_all_particle_thermodynamics with the
 * pre-inverted table (src/sph_workstation.cpp:1730-1734); only the fields
 * fill_thermodynamics writes change.  mode 1: the e -> s inversion of
 * initialize_entropy_and_charge_densities (src/sph_workstation.cpp:290-305)
 * through invert_e_to_s_host and query_thermo_host. */
int orc_preinv_invert(orc_event* ev, int mode)
{
  const orc_preinv* p = ev->eos->pre;
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    if (mode == 0) {
      double o[12];
      preinv_fill(p, FLD(ev, ORC_F_S)[i], FLD(ev, ORC_F_RHOB)[i], o);
      FLD(ev, ORC_F_T)[i] = o[0]; FLD(ev, ORC_F_MUB)[i] = o[1];
      FLD(ev, ORC_F_MUS)[i] = o[2]; FLD(ev, ORC_F_MUQ)[i] = o[3];
      FLD(ev, ORC_F_E)[i] = o[4]; FLD(ev, ORC_F_P)[i] = o[5]; FLD(ev, ORC_F_CS2)[i] = o[6];
      FLD(ev, ORC_F_DWDS)[i] = o[7]; FLD(ev, ORC_F_DWDB)[i] = o[8];
      FLD(ev, ORC_F_DWDS_S)[i] = o[9]; FLD(ev, ORC_F_DWDQ)[i] = o[10];
      FLD(ev, ORC_F_W)[i] = o[11];
    } else {
      const double rhoB = FLD(ev, ORC_F_IN_B)[i];
      const double s = preinv_e_to_s(p, FLD(ev, ORC_F_IN_E)[i], rhoB);
      double o[10];
      preinv_query(p, s, rhoB, o);
      FLD(ev, ORC_F_T)[i] = o[0]; FLD(ev, ORC_F_MUB)[i] = o[1];
      FLD(ev, ORC_F_MUS)[i] = o[2]; FLD(ev, ORC_F_MUQ)[i] = o[3];
      FLD(ev, ORC_F_E)[i] = o[4]; FLD(ev, ORC_F_P)[i] = o[5]; FLD(ev, ORC_F_CS2)[i] = o[6];
      FLD(ev, ORC_F_S)[i] = s;
      FLD(ev, ORC_F_RHOB)[i] = rhoB;
      FLD(ev, ORC_F_RHOS)[i] = 0.0; FLD(ev, ORC_F_RHOQ)[i] = 0.0;
      FLD(ev, ORC_F_W)[i] = o[9];
      FLD(ev, ORC_F_DWDS)[i] = o[7]; FLD(ev, ORC_F_DWDB)[i] = o[8];
      FLD(ev, ORC_F_DWDS_S)[i] = 0.0; FLD(ev, ORC_F_DWDQ)[i] = 0.0;
      FLD(ev, ORC_F_EOS_IS_TABLE)[i] = 1.0;
      FLD(ev, ORC_F_IN_S)[i] = s;
    }
  }
  return ORC_OK;
}

/* point queries for tests: which = 0 fill, 1 query, 2 e -> s (out[0] = s) */
void orc_preinv_point(orc_eos* e, int which, double a, double b, double out[12])
{
  const orc_preinv* p = e->pre;
  if (!p) return;
  if (which == 0) preinv_fill(p, a, b, out);
  else if (which == 1) preinv_query(p, a, b, out);
  else out[0] = preinv_e_to_s(p, a, b);
}
