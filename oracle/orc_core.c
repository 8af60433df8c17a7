/*
 * orc_core.c -- CPU restatement of CCAKE's per-timestep SPH path:
 * kernel, neighbour list, smoothing, gradients, EoS loop, RK2, freeze-out
 * status and reductions (test infrastructure only; see ccake_oracle.h).
 */
#include "orc_internal.h"
#include <string.h>

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_nthreads(void)
{
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}
int orc_get_threads(void) { return orc_nthreads(); }

/* The reference's EoS loop is a serial host loop
 * (src/sph_workstation.cpp:1737-1762); orc_set_serial_eos(1) runs the
 * oracle's EoS inversion on one thread to time that variant.            */
static int g_serial_eos = 0;
void orc_set_serial_eos(int on) { g_serial_eos = on != 0; }

/* Wall time per phase of SPHWorkstation::get_time_derivatives, under the
 * reference's phase names 01-10 (src/sph_workstation.cpp:1543-1583, the
 * CCAKE_PHASE profiler of include/phase_profiler.h), plus the RK update and
 * the freeze-out step of advance_timestep as phases 11 and 12.           */
static double g_phase_s[ORC_NUM_PHASES];
static long long g_phase_n[ORC_NUM_PHASES];
void orc_phase_reset(void) { memset(g_phase_s, 0, sizeof g_phase_s); memset(g_phase_n, 0, sizeof g_phase_n); }
void orc_phase_read(double* seconds, long long* calls)
{
  for (int k = 0; k < ORC_NUM_PHASES; ++k) { seconds[k] = g_phase_s[k]; calls[k] = g_phase_n[k]; }
}
static double now_s(void)
{
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}
#define PHASE(k, stmt) do { const double t_ = now_s(); stmt; g_phase_s[k] += now_s() - t_; g_phase_n[k]++; } while (0)

/* ------------------------------------------------------------------ */
/* SPHkernel<D> (src/kernel.cpp:10-118): cubic spline, q = r/h.        */
static const double KNORM[4]  = { 0.0, 2./3., 10./(7*ORC_PI), 1./ORC_PI };
static const double KNORM2[4] = { 0.0, 1./6., 5./(14.*ORC_PI), 1./(4*ORC_PI) };
static const double KGRAD[4]  = { 0.0, -1./2., -15./(14.*ORC_PI), -3./(4.*ORC_PI) };
static const double KGRAD2[4] = { 0.0, 1., 15./(7.*ORC_PI), 3./(2.*ORC_PI) };

double orc_kernel(int D, double distance, double hT)
{
  const double norm = KNORM[D]/pow(hT, D);
  const double norm2 = KNORM2[D]/pow(hT, D);
  const double q = distance/hT;
  if (q >= 2.0) return 0.0;
  if (q >= 1.0) return norm2*(2.0 - q)*(2.0 - q)*(2.0 - q);
  const double qq = q*q;
  return norm*(1 - 1.5*qq + 0.75*q*qq);
}

void orc_grad_kernel(int D, const double* rel, double r, double h, double* grad)
{
  const double ng = KGRAD[D]/pow(h, D+1);
  const double ng2 = KGRAD2[D]/pow(h, D+1);
  const double q = r/h;
  if (q >= 2.0) { for (int d = 0; d < D; ++d) grad[d] = 0.0; return; }
  if (q >= 1.0) { for (int d = 0; d < D; ++d) grad[d] = (ng/r)*(2.0 - q)*(2.0 - q)*rel[d]; return; }
  for (int d = 0; d < D; ++d) grad[d] = ng2*(1.5*q - 2)*rel[d]/h;
}

static double distance(int D, const double* a, const double* b)
{
  double d = 0.0;
  for (int k = 0; k < D; ++k) { double df = a[k] - b[k]; d += df*df; }
  return sqrt(d);
}

/* The pair's kernel value: SPHkernel<D>::kernel(distance, hT), or for D = 3
 * with the factorized kernel W_2D(r_perp, hT) W_1D(|d eta|, hEta)
 * (src/kernel.cpp:137-145, chosen at src/sph_workstation.cpp:502, 531). */
static double pair_kernel(const orc_params* P, const double* ri, const double* rj)
{
  if (P->dim == 3 && P->factorized_kernel)
    return orc_kernel(2, distance(2, ri, rj), P->hT)*orc_kernel(1, fabs(ri[2] - rj[2]), P->hEta);
  return orc_kernel(P->dim, distance(P->dim, ri, rj), P->hT);
}
static double self_kernel(const orc_params* P)
{
  if (P->dim == 3 && P->factorized_kernel) return orc_kernel(2, 0.0, P->hT)*orc_kernel(1, 0.0, P->hEta);
  return orc_kernel(P->dim, 0.0, P->hT);
}
/* Its gradient with respect to r_a, rel = r_a - r_b: SPHkernel<D>::gradKernel,
 * or the factorized D = 3 gradient (src/kernel.cpp:149-168)
 * (grad_perp W_2D W_1D, W_2D dW_1D/d eta). */
static void pair_grad_kernel(const orc_params* P, const double* rel, const double* pa, const double* pb,
                             double* gK)
{
  if (P->dim == 3 && P->factorized_kernel) {
    const double rp = sqrt(rel[0]*rel[0] + rel[1]*rel[1]);
    const double de = fabs(rel[2]);
    double g2[2], g1;
    orc_grad_kernel(2, rel, rp, P->hT, g2);
    orc_grad_kernel(1, rel + 2, de, P->hEta, &g1);
    const double w2 = orc_kernel(2, rp, P->hT), w1 = orc_kernel(1, de, P->hEta);
    gK[0] = g2[0]*w1; gK[1] = g2[1]*w1; gK[2] = w2*g1;
    return;
  }
  orc_grad_kernel(P->dim, rel, distance(P->dim, pa, pb), P->hT, gK);
}

/* ------------------------------------------------------------------ */
/* freeze_array storage (include/freeze_out.h:29-35), one member per array */
#define SNAP_MEMBERS(X) \
  X(shear[0][0]) X(shear[0][1]) X(shear[0][2]) X(shear[1][0]) X(shear[1][1]) X(shear[1][2]) \
  X(shear[2][0]) X(shear[2][1]) X(shear[2][2]) X(r[0]) X(r[1]) X(u[0]) X(u[1]) \
  X(gradP[0]) X(gradP[1]) X(gradE[0]) X(gradE[1]) X(s) X(e) X(rhoB) X(rhoS) X(rhoQ) \
  X(T) X(muB) X(muS) X(muQ) X(theta) X(bulk) X(sigma_lab) X(shear33) X(shv_nabla_u) X(wfz) X(cs2fz) \
  X(diff[0]) X(diff[1]) X(diff[2]) X(diff[3]) X(diff[4]) X(diff[5]) X(diff[6]) X(diff[7]) \
  X(diff[8]) X(diff[9]) X(diff[10]) X(diff[11])

static orc_snap* snap_alloc(int n)
{
  orc_snap* f = (orc_snap*)calloc(1, sizeof(orc_snap));
#define ALLOC_(m) f->m = (double*)calloc((size_t)n + 1, sizeof(double));
  SNAP_MEMBERS(ALLOC_)
#undef ALLOC_
  return f;
}

static void snap_free(orc_snap* f)
{
  if (!f) return;
#define FREE_(m) free(f->m);
  SNAP_MEMBERS(FREE_)
#undef FREE_
  free(f);
}

static void snap_zero(orc_snap* f, int n)
{
#define ZERO_(m) memset(f->m, 0, (size_t)n * sizeof(double));
  SNAP_MEMBERS(ZERO_)
#undef ZERO_
}

/* FreezeOut::shift_snapshot (include/freeze_out.h:383-419): copy the masked
 * particles of src into dst. */
static void snap_shift(orc_snap* dst, const orc_snap* src, const unsigned char* mask, int n)
{
  for (int i = 0; i < n; ++i) {
    if (!mask[i]) continue;
#define SHIFT_(m) dst->m[i] = src->m[i];
    SNAP_MEMBERS(SHIFT_)
#undef SHIFT_
  }
}

/* The particle's current state into a snapshot (freezeout_first_time_step,
 * freezeout_second_time_step, setup_next_step; include/freeze_out.h:227-318,
 * 594-631), D = 2. */
static void snap_take(orc_snap* f, orc_event* ev)
{
  for (int i = 0; i < ev->n; ++i) {
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) f->shear[a][b][i] = FLD(ev, ORC_F_SHV0 + 4*a + b)[i];
    for (int d = 0; d < 2; ++d) {
      f->r[d][i] = FLD(ev, ORC_F_X + d)[i];
      f->u[d][i] = FLD(ev, ORC_F_UX + d)[i];
      f->gradP[d][i] = FLD(ev, ORC_F_GRADP0 + d)[i];
      f->gradE[d][i] = FLD(ev, ORC_F_GRADE0 + d)[i];
    }
    f->s[i] = FLD(ev, ORC_F_S)[i];       f->e[i] = FLD(ev, ORC_F_E)[i];
    f->rhoB[i] = FLD(ev, ORC_F_RHOB)[i]; f->rhoS[i] = FLD(ev, ORC_F_RHOS)[i]; f->rhoQ[i] = FLD(ev, ORC_F_RHOQ)[i];
    f->T[i] = FLD(ev, ORC_F_T)[i];       f->muB[i] = FLD(ev, ORC_F_MUB)[i];
    f->muS[i] = FLD(ev, ORC_F_MUS)[i];   f->muQ[i] = FLD(ev, ORC_F_MUQ)[i];
    f->theta[i] = FLD(ev, ORC_F_THETA)[i]; f->bulk[i] = FLD(ev, ORC_F_BULK)[i];
    f->sigma_lab[i] = FLD(ev, ORC_F_SIGMA_LAB)[i];
    f->shear33[i] = FLD(ev, ORC_F_SHV0 + 15)[i];
    f->shv_nabla_u[i] = FLD(ev, ORC_F_SHV_NABLA_U)[i];
    f->wfz[i] = FLD(ev, ORC_F_W)[i];     f->cs2fz[i] = FLD(ev, ORC_F_CS2)[i];
    for (int k = 0; k < 12; ++k) f->diff[k][i] = FLD(ev, ORC_F_Q0 + k)[i];
  }
}

orc_event* orc_event_create(const orc_params* p, orc_eos* eos, int n)
{
  orc_event* ev = (orc_event*)calloc(1, sizeof(orc_event));
  ev->p = *p;
  ev->eos = eos;
  ev->n = n;
  ev->t = p->t0;
  for (int f = 0; f < ORC_NUM_FIELDS; ++f) ev->F[f] = (double*)calloc((size_t)n + 1, sizeof(double));
  for (int f = 0; f < ORC_NUM_IFIELDS; ++f) ev->I[f] = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  for (int d = 0; d < 3; ++d) { ev->c_u[d] = calloc((size_t)n+1, 8); ev->c_r[d] = calloc((size_t)n+1, 8); }
  ev->c_s = calloc((size_t)n+1, 8); ev->c_B = calloc((size_t)n+1, 8); ev->c_S = calloc((size_t)n+1, 8);
  ev->c_Q = calloc((size_t)n+1, 8); ev->c_bulk = calloc((size_t)n+1, 8); ev->c_E0 = calloc((size_t)n+1, 8);
  for (int k = 0; k < 5; ++k) ev->c_shv[k] = calloc((size_t)n+1, 8);
  for (int k = 0; k < 9; ++k) ev->c_q[k] = calloc((size_t)n+1, 8);
  for (int k = 0; k < 4; ++k) { ev->fback_e[k] = calloc((size_t)n+1, 8); ev->fback_gradP0[k] = calloc((size_t)n+1, 8); }
  ev->frz1_gradP0 = calloc((size_t)n+1, 8); ev->frz2_gradP0 = calloc((size_t)n+1, 8);
  ev->nb_off = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  if (p->particlization) {
    for (int k = 0; k < 6; ++k) ev->snap[k] = snap_alloc(n);
    ev->fo_rows = (double*)calloc((size_t)n * ORC_FO_ROW + 1, sizeof(double));
  }
  for (int i = 0; i < n; ++i) {
    FLD(ev, ORC_F_MASS_S)[i] = FLD(ev, ORC_F_MASS_B)[i] = 1.0;
    FLD(ev, ORC_F_MASS_SS)[i] = FLD(ev, ORC_F_MASS_Q)[i] = 1.0;
    FLD(ev, ORC_F_GAMMA)[i] = 1.0;
  }
  return ev;
}

void orc_event_destroy(orc_event* ev)
{
  if (!ev) return;
  for (int f = 0; f < ORC_NUM_FIELDS; ++f) free(ev->F[f]);
  for (int f = 0; f < ORC_NUM_IFIELDS; ++f) free(ev->I[f]);
  for (int d = 0; d < 3; ++d) { free(ev->c_u[d]); free(ev->c_r[d]); }
  free(ev->c_s); free(ev->c_B); free(ev->c_S); free(ev->c_Q); free(ev->c_bulk); free(ev->c_E0);
  for (int k = 0; k < 5; ++k) free(ev->c_shv[k]);
  for (int k = 0; k < 9; ++k) free(ev->c_q[k]);
  free(ev->k4);
  for (int k = 0; k < 4; ++k) { free(ev->fback_e[k]); free(ev->fback_gradP0[k]); }
  free(ev->frz1_gradP0); free(ev->frz2_gradP0);
  free(ev->nb_off); free(ev->nb_idx);
  for (int k = 0; k < 6; ++k) snap_free(ev->snap[k]);
  free(ev->fo_rows);
  free(ev);
}

double* orc_field(orc_event* ev, int f) { return (f >= 0 && f < ORC_NUM_FIELDS) ? ev->F[f] : NULL; }
int32_t* orc_ifield(orc_event* ev, int f) { return (f >= 0 && f < ORC_NUM_IFIELDS) ? ev->I[f] : NULL; }
void orc_set_time(orc_event* ev, double t) { ev->t = t; }
double orc_get_time(orc_event* ev) { return ev->t; }

/* ------------------------------------------------------------------ */
/* SystemState::reset_neighbour_list (src/system_state.cpp:650-719).
 * Pair predicate (build contract): j != i and
 *   fl(fl(dx*dx) + fl(dy*dy) [+ fl(dz*dz)]) <= fl(rc*rc),  rc = 2 hT,
 * evaluated without FMA contraction (this file is built with
 * -ffp-contract=off).  Rows hold j ascending.  btrack = row length unless
 * it is the sticky -1 (src/system_state.cpp:712-715).                  */
static int pair_in(int D, const double* xi, const double* xj, double rc2)
{
  double dx = xi[0] - xj[0];
  double r2 = dx*dx;
  if (D > 1) { double dy = xi[1] - xj[1]; r2 = r2 + dy*dy; }
  if (D > 2) { double dz = xi[2] - xj[2]; r2 = r2 + dz*dz; }
  return r2 <= rc2;
}

static int cmp_i32(const void* a, const void* b)
{
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int orc_build_neighbors(orc_event* ev, int brute_force)
{
  const int D = ev->p.dim, n = ev->n;
  const double rc = 2.0*ev->p.hT, rc2 = rc*rc;
  double* X[3] = { FLD(ev, ORC_F_X), FLD(ev, ORC_F_Y), FLD(ev, ORC_F_ETA) };
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  /* uniform cell grid with cell = rc over the current bounding box */
  double lo[3] = {0,0,0}, hi[3] = {0,0,0};
  int nc[3] = {1,1,1};
  for (int d = 0; d < D; ++d) {
    lo[d] = X[d][0]; hi[d] = X[d][0];
    for (int i = 1; i < n; ++i) { if (X[d][i] < lo[d]) lo[d] = X[d][i]; if (X[d][i] > hi[d]) hi[d] = X[d][i]; }
    nc[d] = (int)floor((hi[d] - lo[d])/rc) + 1;
    if (nc[d] < 1) nc[d] = 1;
  }
  const int64_t ncell = (int64_t)nc[0]*nc[1]*nc[2];
  int64_t* head = NULL; int32_t* nxt = NULL; int32_t* cellof = NULL;
  if (!brute_force) {
    head = (int64_t*)malloc((size_t)(ncell + 1)*sizeof(int64_t));
    for (int64_t c = 0; c <= ncell; ++c) head[c] = 0;
    cellof = (int32_t*)malloc((size_t)n*sizeof(int32_t));
    nxt = (int32_t*)malloc((size_t)n*sizeof(int32_t));
    for (int i = 0; i < n; ++i) {
      int c[3] = {0,0,0};
      for (int d = 0; d < D; ++d) { c[d] = (int)floor((X[d][i] - lo[d])/rc); if (c[d] >= nc[d]) c[d] = nc[d]-1; if (c[d] < 0) c[d] = 0; }
      cellof[i] = (int32_t)((c[2]*nc[1] + c[1])*nc[0] + c[0]);
      head[cellof[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) head[c+1] += head[c];
    int64_t* fill = (int64_t*)malloc((size_t)ncell*sizeof(int64_t));
    memcpy(fill, head, (size_t)ncell*sizeof(int64_t));
    for (int i = 0; i < n; ++i) nxt[fill[cellof[i]]++] = i;   /* sorted ids per cell */
    free(fill);
  }
  /* pass 1: counts; pass 2: fill */
  for (int pass = 0; pass < 2; ++pass) {
    #pragma omp parallel for schedule(dynamic, 256) num_threads(orc_nthreads())
    for (int i = 0; i < n; ++i) {
      double xi[3] = { X[0][i], D > 1 ? X[1][i] : 0, D > 2 ? X[2][i] : 0 };
      int64_t k = pass ? ev->nb_off[i] : 0;
      if (brute_force) {
        for (int j = 0; j < n; ++j) {
          if (j == i) continue;
          double xj[3] = { X[0][j], D > 1 ? X[1][j] : 0, D > 2 ? X[2][j] : 0 };
          if (pair_in(D, xi, xj, rc2)) { if (pass) ev->nb_idx[k] = j; ++k; }
        }
      } else {
        int ci[3] = {0,0,0};
        for (int d = 0; d < D; ++d) { ci[d] = (int)floor((xi[d] - lo[d])/rc); if (ci[d] >= nc[d]) ci[d] = nc[d]-1; if (ci[d] < 0) ci[d] = 0; }
        for (int dz = (D > 2 ? -1 : 0); dz <= (D > 2 ? 1 : 0); ++dz)
        for (int dy = (D > 1 ? -1 : 0); dy <= (D > 1 ? 1 : 0); ++dy)
        for (int dxx = -1; dxx <= 1; ++dxx) {
          int c[3] = { ci[0] + dxx, ci[1] + dy, ci[2] + dz };
          if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= nc[0] || c[1] >= nc[1] || c[2] >= nc[2]) continue;
          int64_t cell = ((int64_t)c[2]*nc[1] + c[1])*nc[0] + c[0];
          for (int64_t q = head[cell]; q < head[cell+1]; ++q) {
            int j = nxt[q];
            if (j == i) continue;
            double xj[3] = { X[0][j], D > 1 ? X[1][j] : 0, D > 2 ? X[2][j] : 0 };
            if (pair_in(D, xi, xj, rc2)) { if (pass) ev->nb_idx[k] = j; ++k; }
          }
        }
      }
      if (!pass) cnt[i] = k;
      else qsort(ev->nb_idx + ev->nb_off[i], (size_t)(k - ev->nb_off[i]), sizeof(int32_t), cmp_i32);
    }
    if (!pass) {
      ev->nb_off[0] = 0;
      for (int i = 0; i < n; ++i) ev->nb_off[i+1] = ev->nb_off[i] + cnt[i];
      ev->n_pairs = ev->nb_off[n];
      if (ev->n_pairs > ev->nb_cap) {
        free(ev->nb_idx);
        ev->nb_cap = ev->n_pairs + 1;
        ev->nb_idx = (int32_t*)malloc((size_t)ev->nb_cap*sizeof(int32_t));
      }
    }
  }
  int32_t* bt = ev->I[ORC_I_BTRACK];
  for (int i = 0; i < n; ++i) bt[i] = (bt[i] == -1) ? -1 : (int32_t)cnt[i];
  free(cnt); free(head); free(nxt); free(cellof);
  return ORC_OK;
}

int64_t orc_pair_count(orc_event* ev) { return ev->n_pairs; }

int orc_export_pairs(orc_event* ev, int32_t* pi, int32_t* pj, int64_t cap)
{
  if (ev->n_pairs > cap) return 7;
  for (int i = 0; i < ev->n; ++i)
    for (int64_t k = ev->nb_off[i]; k < ev->nb_off[i+1]; ++k) { pi[k] = i; pj[k] = ev->nb_idx[k]; }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* EoM_default::update_velocity (src/eom_default.cpp:17-38) */
void orc_gamma_velocity(orc_event* ev)
{
  const int D = ev->p.dim;
  const double t2 = ev->t*ev->t;
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    double u[3] = {0,0,0}, uc[3];
    for (int d = 0; d < D; ++d) u[d] = uc[d] = FLD(ev, ORC_F_UX + d)[i];
    if (D == 2) { uc[0] = -uc[0]; uc[1] = -uc[1]; }
    else { for (int d = 0; d < D-1; ++d) uc[d] = -uc[d]; uc[D-1] *= -t2; }
    double c = 0.0;
    for (int d = 0; d < D; ++d) c += uc[d]*u[d];
    const double g = sqrt(1 - c);
    FLD(ev, ORC_F_GAMMA)[i] = g;
    for (int d = 0; d < D; ++d) FLD(ev, ORC_F_VX + d)[i] = u[d]/g;
  }
}

static void pos_of(orc_event* ev, int i, double* r)
{
  r[0] = FLD(ev, ORC_F_X)[i]; r[1] = FLD(ev, ORC_F_Y)[i]; r[2] = FLD(ev, ORC_F_ETA)[i];
}

/* SPHWorkstation::smooth_all_particle_fields (src/sph_workstation.cpp:494-578) */
void orc_smooth_fields(orc_event* ev)
{
  const double t = ev->t;
  const double kern0 = self_kernel(&ev->p);
  double* ms = FLD(ev, ORC_F_MASS_S); double* mB = FLD(ev, ORC_F_MASS_B);
  double* mS = FLD(ev, ORC_F_MASS_SS); double* mQ = FLD(ev, ORC_F_MASS_Q);
  double* es = FLD(ev, ORC_F_EXT_S); double* eB = FLD(ev, ORC_F_EXT_B);
  double* eS = FLD(ev, ORC_F_EXT_SS); double* eQ = FLD(ev, ORC_F_EXT_Q);
  #pragma omp parallel for schedule(dynamic, 256) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    double ss = ms[i]*es[i]*kern0, sB = mB[i]*eB[i]*kern0;
    double sQ = mQ[i]*eQ[i]*kern0, sS = mS[i]*eS[i]*kern0;
    double sl = ms[i]*kern0;
    double ri[3], rj[3];
    pos_of(ev, i, ri);
    for (int64_t k = ev->nb_off[i]; k < ev->nb_off[i+1]; ++k) {
      const int j = ev->nb_idx[k];
      pos_of(ev, j, rj);
      const double kern = pair_kernel(&ev->p, ri, rj);
      sl += ms[j]*kern;
      ss += ms[j]*es[j]*kern;
      sB += mB[j]*eB[j]*kern;
      sS += mS[j]*eS[j]*kern;
      sQ += mQ[j]*eQ[j]*kern;
    }
    FLD(ev, ORC_F_SM_S)[i] = ss; FLD(ev, ORC_F_SM_B)[i] = sB;
    FLD(ev, ORC_F_SM_SS)[i] = sS; FLD(ev, ORC_F_SM_Q)[i] = sQ;
    FLD(ev, ORC_F_SIGMA_LAB)[i] = sl;
    /* get_LRF = lab/gamma/t (src/eom_default.cpp:198-203) */
    const double g = FLD(ev, ORC_F_GAMMA)[i];
    FLD(ev, ORC_F_SIGMA)[i] = sl/g/t;
    FLD(ev, ORC_F_S)[i] = ss/g/t;
    FLD(ev, ORC_F_RHOB)[i] = sB/g/t;
    FLD(ev, ORC_F_RHOS)[i] = sS/g/t;
    FLD(ev, ORC_F_RHOQ)[i] = sQ/g/t;
  }
}

/* SPHWorkstation::calculate_intial_sigma (src/sph_workstation.cpp:582-640) */
void orc_initial_sigma(orc_event* ev)
{
  const double t = ev->t;
  const double kern0 = self_kernel(&ev->p);
  double* ms = FLD(ev, ORC_F_MASS_S);
  #pragma omp parallel for schedule(dynamic, 256) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    double sl = ms[i]*kern0, ri[3], rj[3];
    pos_of(ev, i, ri);
    for (int64_t k = ev->nb_off[i]; k < ev->nb_off[i+1]; ++k) {
      const int j = ev->nb_idx[k];
      pos_of(ev, j, rj);
      sl += ms[j]*pair_kernel(&ev->p, ri, rj);
    }
    FLD(ev, ORC_F_SIGMA_LAB)[i] = sl;
    FLD(ev, ORC_F_SIGMA)[i] = sl/FLD(ev, ORC_F_GAMMA)[i]/t;
  }
}

/* SPHWorkstation::smooth_all_particle_gradients (src/sph_workstation.cpp:796-986),
 * diffusion terms off.  Owner a, neighbour b. */
void orc_smooth_gradients(orc_event* ev)
{
  const int D = ev->p.dim;
  const double hT = ev->p.hT;
  const int using_shear = ev->p.using_shear;
  #pragma omp parallel for schedule(dynamic, 256) num_threads(orc_nthreads())
  for (int a = 0; a < ev->n; ++a) {
    double gP[3] = {0,0,0}, gE[3] = {0,0,0}, gB[3] = {0,0,0}, gS[3] = {0,0,0}, dS[3] = {0,0,0};
    double gV[3][3] = {{0}};
    double pa[3], pb[3];
    pos_of(ev, a, pa);
    const double p_a = FLD(ev, ORC_F_P)[a], e_a = FLD(ev, ORC_F_E)[a];
    const double sl_a = FLD(ev, ORC_F_SIGMA_LAB)[a];
    const double sigsqra = 1.0/(sl_a*sl_a);
    const double bulk_a = FLD(ev, ORC_F_BULK)[a];
    double va[3] = {0,0,0};
    for (int d = 0; d < D; ++d) va[d] = FLD(ev, ORC_F_VX + d)[a];
    double s0a[3] = {0,0,0}, mina[3][3] = {{0}};
    for (int d = 0; d < D; ++d) {
      s0a[d] = FLD(ev, ORC_F_SHV0 + 0*4 + d + 1)[a];
      for (int e2 = 0; e2 < D; ++e2) mina[d][e2] = FLD(ev, ORC_F_SHV0 + (d+1)*4 + e2 + 1)[a];
    }
    /* diffusion sums (src/sph_workstation.cpp:909-967): grad q^0, div q, grad alpha */
    const int using_diff = ev->p.using_diffusion;
    double gqa[3] = {0,0,0}, dqa[3] = {0,0,0}, gal[3][3] = {{0}};
    double al_a[3] = {0,0,0}, q0_a[3] = {0,0,0}, qv_a[3][3] = {{0}};
    if (using_diff) {
      const double Ta = FLD(ev, ORC_F_T)[a];
      al_a[0] = FLD(ev, ORC_F_MUB)[a]/Ta; al_a[1] = FLD(ev, ORC_F_MUS)[a]/Ta; al_a[2] = FLD(ev, ORC_F_MUQ)[a]/Ta;
      for (int c = 0; c < 3; ++c) {
        q0_a[c] = FLD(ev, ORC_F_Q0 + 4*c)[a];
        for (int d = 0; d < D; ++d) qv_a[c][d] = FLD(ev, ORC_F_Q0 + 4*c + d + 1)[a];
      }
    }
    for (int64_t k = ev->nb_off[a]; k < ev->nb_off[a+1]; ++k) {
      const int b = ev->nb_idx[k];
      pos_of(ev, b, pb);
      double rel[3], gK[3];
      for (int d = 0; d < D; ++d) rel[d] = pa[d] - pb[d];
      pair_grad_kernel(&ev->p, rel, pa, pb, gK);
      const double p_b = FLD(ev, ORC_F_P)[b], e_b = FLD(ev, ORC_F_E)[b];
      const double sl_b = FLD(ev, ORC_F_SIGMA_LAB)[b];
      const double sigsqrb = 1.0/(sl_b*sl_b);
      const double m_b = FLD(ev, ORC_F_MASS_S)[b];
      const double bulk_b = FLD(ev, ORC_F_BULK)[b];
      double vb[3] = {0,0,0};
      for (int d = 0; d < D; ++d) vb[d] = FLD(ev, ORC_F_VX + d)[b];
      double ssK[3];
      const double msl = m_b*sl_a;
      for (int d = 0; d < D; ++d) ssK[d] = msl*gK[d];
      if (using_diff) {
        const double Tb = FLD(ev, ORC_F_T)[b];
        const double al_b[3] = { FLD(ev, ORC_F_MUB)[b]/Tb, FLD(ev, ORC_F_MUS)[b]/Tb, FLD(ev, ORC_F_MUQ)[b]/Tb };
        double cva = 0.0;
        for (int d = 0; d < D; ++d) cva += ssK[d]*va[d];
        for (int c = 0; c < 3; ++c) {
          gqa[c] += cva*(sigsqrb*FLD(ev, ORC_F_Q0 + 4*c)[b] + sigsqra*q0_a[c]);
          double dv = 0.0;
          for (int d = 0; d < D; ++d)
            dv += sigsqrb*ssK[d]*FLD(ev, ORC_F_Q0 + 4*c + d + 1)[b] + sigsqra*ssK[d]*qv_a[c][d];
          dqa[c] += dv;
          for (int d = 0; d < D; ++d) gal[d][c] += (m_b/sl_a)*(al_b[c] - al_a[c])*gK[d];
        }
      }
      const double cP = sigsqrb*p_b + sigsqra*p_a;
      const double cE = (m_b/sl_a)*(e_b - e_a);
      const double cB = bulk_b/sl_b/sl_b + bulk_a/sl_a/sl_a;
      for (int d = 0; d < D; ++d) {
        gP[d] += cP*ssK[d];
        gE[d] += cE*gK[d];
        gB[d] += cB*ssK[d];
        for (int e2 = 0; e2 < D; ++e2) gV[d][e2] += (m_b/sl_a)*(vb[e2] - va[e2])*gK[d];
      }
      if (using_shear && D > 1) {
        double cv = 0.0;
        for (int d = 0; d < D; ++d) cv += ssK[d]*va[d];
        double s0b[3];
        for (int d = 0; d < D; ++d) s0b[d] = FLD(ev, ORC_F_SHV0 + d + 1)[b];
        for (int d = 0; d < D; ++d) gS[d] += cv*(sigsqrb*s0b[d] + sigsqra*s0a[d]);
        /* sigsqrb*sigsigK*transpose(vminib) + sigsqra*sigsigK*transpose(vminia) */
        for (int j = 0; j < D; ++j) {
          double tb = 0.0, ta = 0.0;
          for (int ii = 0; ii < D; ++ii) {
            tb += (sigsqrb*ssK[ii])*FLD(ev, ORC_F_SHV0 + (j+1)*4 + ii + 1)[b];
            ta += (sigsqra*ssK[ii])*mina[j][ii];
          }
          dS[j] += tb + ta;
        }
      }
    }
    if (using_diff) {
      for (int c = 0; c < 3; ++c) {
        FLD(ev, ORC_F_GRAD_QA0 + c)[a] = gqa[c];
        FLD(ev, ORC_F_DIV_QA0 + c)[a] = dqa[c];
        for (int d = 0; d < D; ++d) FLD(ev, ORC_F_GRAD_ALPHA0 + 3*d + c)[a] = gal[d][c];
      }
    }
    for (int d = 0; d < 3; ++d) {
      FLD(ev, ORC_F_GRADP0 + d)[a] = gP[d];
      FLD(ev, ORC_F_GRADE0 + d)[a] = gE[d];
      FLD(ev, ORC_F_GRADBULK0 + d)[a] = gB[d];
      FLD(ev, ORC_F_GRADSHEAR0 + d)[a] = gS[d];
      FLD(ev, ORC_F_DIVSHEAR0 + d)[a] = dS[d];
      for (int e2 = 0; e2 < 3; ++e2) FLD(ev, ORC_F_GRADV0 + 3*d + e2)[a] = gV[d][e2];
    }
  }
}

/* ------------------------------------------------------------------ */
/* SPHWorkstation::update_all_particle_thermodynamics (1713-1772), mode 0,
 * and the e-inversion of initialize_entropy_and_charge_densities via
 * locate_phase_diagram_point_eBSQ (1791-1851), mode 1.  Returns the
 * first failure status (EoS_NO_SOLUTION) or OK.                         */
int orc_eos_invert(orc_event* ev, int mode)
{
  if (ev->eos->pre) return orc_preinv_invert(ev, mode);   /* online_inverter_enabled = false */
  int fail = 0;
  #pragma omp parallel for schedule(dynamic, 64) num_threads(g_serial_eos ? 1 : orc_nthreads()) reduction(|:fail)
  for (int i = 0; i < ev->n; ++i) {
    double x[4] = { FLD(ev, ORC_F_T)[i], FLD(ev, ORC_F_MUB)[i], FLD(ev, ORC_F_MUQ)[i], FLD(ev, ORC_F_MUS)[i] };
    int member = ev->I[ORC_I_EOS_ID][i];
    double th[31];
    double tgt[4];
    if (mode == 0) {
      tgt[0] = FLD(ev, ORC_F_S)[i]; tgt[1] = FLD(ev, ORC_F_RHOB)[i];
      tgt[2] = FLD(ev, ORC_F_RHOS)[i]; tgt[3] = FLD(ev, ORC_F_RHOQ)[i];
      if (!orc_eos_invert_point(ev->eos, 0, tgt, x, &member, th)) { fail = 1; continue; }
    } else {
      tgt[0] = FLD(ev, ORC_F_IN_E)[i]; tgt[1] = FLD(ev, ORC_F_IN_B)[i];
      tgt[2] = FLD(ev, ORC_F_IN_SS)[i]; tgt[3] = FLD(ev, ORC_F_IN_Q)[i];
      if (!orc_eos_invert_point(ev->eos, 1, tgt, x, &member, th)) {
        /* freeze the particle and retry at zero charge (1825-1848) */
        ev->I[ORC_I_FREEZE][i] = 4;
        double z[4] = { tgt[0], 0.0, 0.0, 0.0 };
        if (!orc_eos_invert_point(ev->eos, 1, z, x, &member, th)) { fail = 1; continue; }
      }
      FLD(ev, ORC_F_IN_S)[i] = th[5];
    }
    for (int k = 0; k < 31; ++k) FLD(ev, ORC_F_T + k)[i] = th[k];
    ev->I[ORC_I_EOS_ID][i] = member;
  }
  return fail ? ORC_ERR_EOS_NO_SOLUTION : ORC_OK;
}

/* ------------------------------------------------------------------ */
/* SPHWorkstation::get_time_derivatives (src/sph_workstation.cpp:1509-1612) */
int orc_get_time_derivatives(orc_event* ev)
{
  int st = 0;
  PHASE(1, orc_build_neighbors(ev, 0));    /* 01_neighbour_list */
  PHASE(2, orc_gamma_velocity(ev));        /* 02_gamma_velocity */
  PHASE(3, orc_smooth_fields(ev));         /* 03_smooth_fields */
  PHASE(4, st = orc_eos_invert(ev, 0));    /* 04_thermodynamics */
  if (st) return st;
  PHASE(5, orc_reset_pi_tensor(ev));       /* 05_reset_pi_tensor */
  if (ev->p.using_diffusion && ev->p.diffusion_regulator)
    PHASE(6, orc_regulate_diffusion(ev));  /* 06_regulate_diffusion */
  PHASE(8, orc_viscosities(ev));           /* 08_viscosities */
  PHASE(9, orc_smooth_gradients(ev));      /* 09_smooth_gradients */
  PHASE(10, orc_eom(ev));                  /* 10_time_derivatives */
  return ORC_OK;
}

/* Evolver::set_current_timestep_quantities (src/evolver.cpp:124-168) */
void orc_cache_state(orc_event* ev)
{
  for (int i = 0; i < ev->n; ++i) {
    for (int k = 0; k < 5; ++k) ev->c_shv[k][i] = FLD(ev, ORC_F_EXT_SHV0 + k)[i];
    for (int k = 0; k < 9; ++k) ev->c_q[k][i] = FLD(ev, ORC_F_EXT_Q0 + k)[i];
    for (int d = 0; d < ev->p.dim; ++d) { ev->c_u[d][i] = FLD(ev, ORC_F_UX + d)[i]; ev->c_r[d][i] = FLD(ev, ORC_F_X + d)[i]; }
    ev->c_s[i] = FLD(ev, ORC_F_EXT_S)[i];
    ev->c_bulk[i] = FLD(ev, ORC_F_EXT_BULK)[i];
    ev->c_B[i] = FLD(ev, ORC_F_EXT_B)[i];
    ev->c_S[i] = FLD(ev, ORC_F_EXT_SS)[i];
    ev->c_Q[i] = FLD(ev, ORC_F_EXT_Q)[i];
    ev->c_E0[i] = FLD(ev, ORC_F_EZ)[i];
  }
}

/* Evolver::step_rk update (src/evolver.cpp:363-484) */
int orc_rk_update(orc_event* ev, double dt, double t0)
{
  const int D = ev->p.dim;
  int fail = 0;
  for (int i = 0; i < ev->n; ++i) {
    double es = ev->c_s[i] + dt*FLD(ev, ORC_F_DEXT_S)[i];
    const double eb = ev->c_bulk[i] + dt*FLD(ev, ORC_F_DEXT_BULK)[i];
    const double rB = ev->c_B[i] + dt*FLD(ev, ORC_F_DEXT_B)[i];
    const double rS = ev->c_S[i] + dt*FLD(ev, ORC_F_DEXT_SS)[i];
    const double rQ = ev->c_Q[i] + dt*FLD(ev, ORC_F_DEXT_Q)[i];
    const double ez = ev->c_E0[i] + dt*FLD(ev, ORC_F_DEZ)[i];
    const int fr = ev->I[ORC_I_FREEZE][i];
    if (es < 0.0 && fr > 3) es = 1.e-3;
    else if (es < 0.0) fail = 1;
    FLD(ev, ORC_F_EXT_S)[i] = es;
    FLD(ev, ORC_F_EXT_BULK)[i] = eb;
    FLD(ev, ORC_F_EXT_B)[i] = rB;
    FLD(ev, ORC_F_EXT_SS)[i] = rS;
    FLD(ev, ORC_F_EXT_Q)[i] = rQ;
    FLD(ev, ORC_F_EZ)[i] = ez;
    for (int d = 0; d < D; ++d) {
      FLD(ev, ORC_F_X + d)[i] = ev->c_r[d][i] + dt*FLD(ev, ORC_F_VX + d)[i];
      FLD(ev, ORC_F_UX + d)[i] = ev->c_u[d][i] + dt*FLD(ev, ORC_F_DUDT0 + d)[i];
    }
    for (int k = 0; k < 5; ++k)
      FLD(ev, ORC_F_EXT_SHV0 + k)[i] = ev->c_shv[k][i] + dt*FLD(ev, ORC_F_DEXT_SHV0 + k)[i];
    if (ev->p.using_diffusion)
      for (int k = 0; k < 9; ++k) FLD(ev, ORC_F_EXT_Q0 + k)[i] = ev->c_q[k][i] + dt*FLD(ev, ORC_F_DEXT_Q0 + k)[i];
    for (int d = D; d < 3; ++d) { FLD(ev, ORC_F_X + d)[i] = 0.0; FLD(ev, ORC_F_UX + d)[i] = 0.0; }
  }
  ev->t = t0 + dt;
  return fail ? ORC_ERR_NEG_ENTROPY : ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Freeze-out status (include/freeze_out.h:322-626): the status machine of
 * check_freeze_out_status plus the energy / gradP snapshot shifts of
 * bsqsvfreezeout and freeze_out_evo that decide it.  Status 3 -> 4 is the
 * hand-over to the hypersurface interpolation (bsqsvinterpolate).       */
static int check_status(orc_event* ev)
{
  const double efc = ev->p.e_freeze;
  int count = 0;
  int32_t* fz = ev->I[ORC_I_FREEZE];
  int32_t* bt = ev->I[ORC_I_BTRACK];
  for (int i = 0; i < ev->n; ++i) {
    const double e = FLD(ev, ORC_F_E)[i];
    switch (fz[i]) {
      case 0: if (e <= efc) fz[i] = 1; break;
      case 1:
        if (bt[i] == -1) fz[i] = 3;
        else if (e > FLD(ev, ORC_F_FRZ1_E)[i]) fz[i] = 1;
        else if (e <= efc) fz[i] = 3;
        else fz[i] = 0;
        break;
      default: break;
    }
  }
  for (int i = 0; i < ev->n; ++i) count += (fz[i] == 3);
  ev->number_part_fo += count;
  return count;
}

static void snap(double* de, double* dg, const double* se, const double* sg, const unsigned char* mask, int n)
{
  for (int i = 0; i < n; ++i) if (mask[i]) { de[i] = se[i]; dg[i] = sg[i]; }
}

/* FreezeOut::bsqsvinterpolate (include/freeze_out.h:638-930) for the
 * particles at status 3, D = 2, hyperbolic coordinates: one row each,
 * appended in particle order with the column order of
 * Output<2,EoM_default>::print_freeze_out (src/output.cpp:502-560). */
static void interpolate(orc_event* ev)
{
  const double efc = ev->p.e_freeze;
  const double hbarc_GeVfm = 0.1973269804;            /* include/constants.h:15 */
  int32_t* fz = ev->I[ORC_I_FREEZE];
  int32_t* bt = ev->I[ORC_I_BTRACK];
  const orc_snap* f1 = ev->snap[0];
  const orc_snap* f2 = ev->snap[1];
  for (int i = 0; i < ev->n; ++i) {
    if (fz[i] != 3) continue;
    /* the snapshot closer to the freeze-out energy */
    const int swit = (fabs(f1->e[i] - efc) < fabs(f2->e[i] - efc)) ? 1 : 2;
    const orc_snap* f = (swit == 1) ? f1 : f2;
    const double tau_s = (swit == 1) ? ev->taup : ev->taupp;
    const double tlist = (bt[i] != -1) ? tau_s : tau_s - ev->p.dt;
    double uout[2], gradPsub[2], gradEsub[2], uout_cov[2], gradPsub_contra[2], gradEsub_contra[2];
    for (int d = 0; d < 2; ++d) {
      uout[d] = f->u[d][i]; gradPsub[d] = f->gradP[d][i]; gradEsub[d] = f->gradE[d][i];
      /* milne::Vector<double,2>::make_covariant / make_contravariant
       * (include/milne.hpp:353-364): a sign flip in 2+1D */
      uout_cov[d] = -f->u[d][i];
      gradPsub_contra[d] = f->gradP[d][i] / -1.0;
      gradEsub_contra[d] = f->gradE[d][i] / -1.0;
    }
    double norm2 = 0.0;
    for (int d = 0; d < 2; ++d) norm2 += uout[d]*uout_cov[d];
    const double gsub = sqrt(-norm2 + 1);
    double sigsub = f->sigma_lab[i];
    sigsub /= gsub*tlist;                                  /* "hyperbolic" */
    const double swsub = FLD(ev, ORC_F_MASS_S)[i]/sigsub;
    double sFO = f->s[i];
    double divT[2], divP[2], divE[2], divT_contra[2];
    for (int d = 0; d < 2; ++d) {
      divT[d] = (1.0/sFO)*gradPsub[d];
      divP[d] = gradPsub[d];
      divE[d] = gradEsub[d];
      divT_contra[d] = (1.0/sFO)*gradPsub_contra[d];
    }
    (void)divP;
    const double cs2 = f->cs2fz[i], w = f->wfz[i], bulk = f->bulk[i], theta = f->theta[i];
    const double shv_nabla_u = f->shv_nabla_u[i];
    double inner = 0.0;
    for (int d = 0; d < 2; ++d) inner += uout[d]*gradPsub[d];
    const double divPpress = -(1.0/gsub)*(cs2*(w + bulk)*theta - cs2*shv_nabla_u + inner);   /* arXiv:1406.3333 (C3) */
    double divTtemp = divPpress/sFO;
    inner = 0.0;
    for (int d = 0; d < 2; ++d) inner += uout[d]*gradEsub[d];
    double divEener = -(1.0/gsub)*((w + bulk)*theta - shv_nabla_u + inner);                /* (C2) */
    norm2 = 0.0;
    for (int d = 0; d < 2; ++d) norm2 += divT[d]*divT_contra[d];
    const double insub = divTtemp*divTtemp + norm2;
    double norm = -sqrt(fabs(insub));
    norm = fabs(norm) > 1.E-14 ? norm : 1.E-14;
    divTtemp /= norm;
    for (int d = 0; d < 2; ++d) divT[d] /= norm;
    norm2 = 0.0;
    for (int d = 0; d < 2; ++d) norm2 += divE[d]*gradEsub_contra[d];
    const double insubE = divEener*divEener + norm2;
    double normE = -sqrt(fabs(insubE));
    normE = fabs(normE) > 1.E-14 ? normE : 1.E-14;
    divEener /= normE;
    for (int d = 0; d < 2; ++d) divE[d] /= normE;
    double Tfluc = f->T[i];
    sFO *= pow(Tfluc*hbarc_GeVfm, 3);
    Tfluc *= hbarc_GeVfm;
    double* row = ev->fo_rows + (size_t)ev->fo_count*ORC_FO_ROW;
    double* c = row + 2;
    row[0] = i; row[1] = ev->fo_step;
    c[0] = divEener; c[1] = divE[0]; c[2] = divE[1];
    c[3] = gsub; c[4] = uout[0]; c[5] = uout[1];
    c[6] = swsub; c[7] = bulk;
    c[8] = f->shear[0][0][i]; c[9] = f->shear[1][1][i]; c[10] = f->shear[2][2][i];
    c[11] = f->shear33[i]; c[12] = f->shear[1][2][i];
    c[13] = tlist; c[14] = f->r[0][i]; c[15] = f->r[1][i];
    c[16] = sFO; c[17] = f->e[i]; c[18] = Tfluc;
    c[19] = f->muB[i]; c[20] = f->muS[i]; c[21] = f->muQ[i];
    c[22] = w; c[23] = cs2;
    c[24] = f->rhoB[i]; c[25] = f->rhoS[i]; c[26] = f->rhoQ[i];
    for (int q = 0; q < 12; ++q) c[27 + q] = f->diff[q][i];   /* diffout(charge, mu) */
    c[39] = FLD(ev, ORC_F_EOS_IS_TABLE)[i];
    ev->fo_count++;
    fz[i] = 4;
  }
}

int orc_freeze_out_surface(orc_event* ev, double* rows, int cap)
{
  if (!ev->fo_rows) return 0;
  const int k = ev->fo_count < cap ? ev->fo_count : cap;
  if (rows && k > 0) memcpy(rows, ev->fo_rows, (size_t)k*ORC_FO_ROW*sizeof(double));
  return ev->fo_count;
}

void orc_freeze_out_particles(orc_event* ev)
{
  const int n = ev->n;
  const int curfrz = check_status(ev);
  const int full = ev->fo_rows != NULL;
  orc_snap** S = ev->snap;        /* 0 frz1, 1 frz2, 2 fback, 3 fback2, 4 fback3, 5 fback4 */
  int32_t* fz = ev->I[ORC_I_FREEZE];
  int32_t* bt = ev->I[ORC_I_BTRACK];
  double* e1 = FLD(ev, ORC_F_FRZ1_E); double* e2 = FLD(ev, ORC_F_FRZ2_E);
  double* g1 = ev->frz1_gradP0; double* g2 = ev->frz2_gradP0;
  double** fe = ev->fback_e; double** fg = ev->fback_gradP0;
  ev->fo_step++;
  if (ev->frzc == 0) {                                  /* freezeout_first_time_step */
    ev->taupp = ev->t;
    for (int i = 0; i < n; ++i) { e2[i] = FLD(ev, ORC_F_E)[i]; g2[i] = FLD(ev, ORC_F_GRADP0)[i]; }
    if (full) snap_take(S[1], ev);
    ev->frzc = 1;
    return;
  }
  unsigned char* m = (unsigned char*)calloc((size_t)n, 1);
  if (ev->frzc == 1) {                                  /* freezeout_second_time_step */
    ev->taup = ev->t;
    for (int i = 0; i < n; ++i) { e1[i] = FLD(ev, ORC_F_E)[i]; g1[i] = FLD(ev, ORC_F_GRADP0)[i]; }
    if (full) snap_take(S[0], ev);
    ev->frzc = 2;
    if (curfrz > 0) {
      if (full) interpolate(ev);
      else for (int i = 0; i < n; ++i) if (fz[i] == 3) fz[i] = 4;
    }
    free(m);
    return;
  }
  /* freeze_out_evo */
  for (int i = 0; i < n; ++i) m[i] = (fz[i] < 4 && bt[i] <= 3 && bt[i] > 0);
  snap(fe[3], fg[3], fe[1], fg[1], m, n);     /* fback4 <- fback2 */
  snap(fe[2], fg[2], fe[0], fg[0], m, n);     /* fback3 <- fback  */
  snap(fe[1], fg[1], e2, g2, m, n);           /* fback2 <- frz2   */
  snap(fe[0], fg[0], e1, g1, m, n);           /* fback  <- frz1   */
  if (full) { snap_shift(S[5], S[3], m, n); snap_shift(S[4], S[2], m, n);
              snap_shift(S[3], S[1], m, n); snap_shift(S[2], S[0], m, n); }
  for (int i = 0; i < n; ++i) m[i] = (fz[i] < 4 && bt[i] == 0 && fg[0][i] != 0);
  snap(e2, g2, fe[1], fg[1], m, n);
  snap(e1, g1, fe[0], fg[0], m, n);
  if (full) { snap_shift(S[1], S[3], m, n); snap_shift(S[0], S[2], m, n); }
  for (int i = 0; i < n; ++i) m[i] = (fz[i] < 4 && bt[i] == 0 && fg[0][i] == 0);
  snap(e2, g2, fe[3], fg[3], m, n);
  snap(e1, g1, fe[2], fg[2], m, n);
  if (full) { snap_shift(S[1], S[5], m, n); snap_shift(S[0], S[4], m, n); }
  for (int i = 0; i < n; ++i) if (fz[i] < 4 && bt[i] == 0) { fz[i] = 4; bt[i] = -1; }
  const double tau = ev->t;
  if (curfrz > 0) {
    if (full) interpolate(ev);
    else for (int i = 0; i < n; ++i) if (fz[i] == 3) fz[i] = 4;
  }
  for (int i = 0; i < n; ++i) m[i] = (fz[i] < 4);
  snap(e2, g2, e1, g1, m, n);
  if (full) snap_shift(S[1], S[0], m, n);
  for (int i = 0; i < n; ++i) { e1[i] = FLD(ev, ORC_F_E)[i]; g1[i] = FLD(ev, ORC_F_GRADP0)[i]; }
  if (full) snap_take(S[0], ev);
  ev->taupp = ev->taup;
  ev->taup = tau;
  free(m);
}

/* SPHWorkstation::continue_evolution (src/sph_workstation.cpp:1992-2024) */
int orc_continue_evolution(orc_event* ev, double max_tau)
{
  int keep = 1;
  if (ev->p.particlization) {
    int frozen = 0;
    for (int i = 0; i < ev->n; ++i) frozen += (ev->I[ORC_I_FREEZE][i] >= 4);
    ev->number_part_fo = frozen;
    keep = frozen < ev->n;
  }
  return keep && (ev->t <= max_tau);
}

/* Evolver::advance_timestep_rk2 (src/evolver.cpp:493-517) inside
 * SPHWorkstation::advance_timestep (src/sph_workstation.cpp:2042-2073). */
/* ------------------------------------------------------------------ */
/* The reference's fourth-order step, Evolver::advance_timestep_rk4
 * (src/evolver.cpp:519-561) with update_k (171-212) and update_rk4
 * (215-341), restated as it is (it is not the classical method):
 *   y1 = y0 + dt/2 f(y0, t0)                     (step 1, f not kept)
 *   k2 = f(y1, t0+dt/2);  y2 = y0 + dt/2 k2;     k += k2 / 6
 *   k3 = f(y2, t0+dt/2);  y3 = y0 + dt/2 k3;     k += 2 k3 / 6
 *   k4 = f(y3, t0+dt/2);  y4 = y0 + dt k4, t = t0 + dt;  k += 2 k4 / 6
 *   k5 = f(y4, t0+dt);                           k += 2 k5 / 6
 *   y = y0 + dt k   (weights summing to 7/6), with the negative-entropy rule
 * and the extensive bulk pressure written into the bulk field too (218-341).
 * k of the position is accumulated from v (dr/dt).                       */
enum { K4_U = 0, K4_R = 3, K4_S = 6, K4_B, K4_SS, K4_Q, K4_BULK, K4_E0, K4_SHV, K4_DQ = K4_SHV + 5, K4_N = K4_DQ + 9 };

static double* k4col(orc_event* ev, int c) { return ev->k4 + (size_t)c * ((size_t)ev->n + 1); }

static void rk4_accumulate(orc_event* ev, double w)
{
  const int D = ev->p.dim;
  for (int i = 0; i < ev->n; ++i) {
    for (int k = 0; k < 5; ++k) k4col(ev, K4_SHV + k)[i] += w*FLD(ev, ORC_F_DEXT_SHV0 + k)[i];
    for (int k = 0; k < 9; ++k) k4col(ev, K4_DQ + k)[i] += w*FLD(ev, ORC_F_DEXT_Q0 + k)[i];
    for (int d = 0; d < D; ++d) {
      k4col(ev, K4_U + d)[i] += w*FLD(ev, ORC_F_DUDT0 + d)[i];
      k4col(ev, K4_R + d)[i] += w*FLD(ev, ORC_F_VX + d)[i];
    }
    k4col(ev, K4_B)[i] += w*FLD(ev, ORC_F_DEXT_B)[i];
    k4col(ev, K4_SS)[i] += w*FLD(ev, ORC_F_DEXT_SS)[i];
    k4col(ev, K4_Q)[i] += w*FLD(ev, ORC_F_DEXT_Q)[i];
    k4col(ev, K4_S)[i] += w*FLD(ev, ORC_F_DEXT_S)[i];
    k4col(ev, K4_BULK)[i] += w*FLD(ev, ORC_F_DEXT_BULK)[i];
    k4col(ev, K4_E0)[i] += w*FLD(ev, ORC_F_DEZ)[i];
  }
}

static int rk4_update(orc_event* ev, double dt)
{
  const int D = ev->p.dim;
  int fail = 0;
  for (int i = 0; i < ev->n; ++i) {
    double es = ev->c_s[i] + dt*k4col(ev, K4_S)[i];
    const double rB = ev->c_B[i] + dt*k4col(ev, K4_B)[i];
    const double rS = ev->c_S[i] + dt*k4col(ev, K4_SS)[i];
    const double rQ = ev->c_Q[i] + dt*k4col(ev, K4_Q)[i];
    const double bulk = ev->c_bulk[i] + dt*k4col(ev, K4_BULK)[i];
    const double ez = ev->c_E0[i] + dt*k4col(ev, K4_E0)[i];
    const int fr = ev->I[ORC_I_FREEZE][i];
    if (es < 0.0 && fr > 3) es = 1.e-3;
    else if (es < 0.0) fail = 1;
    FLD(ev, ORC_F_EXT_S)[i] = es;
    FLD(ev, ORC_F_EXT_BULK)[i] = bulk;
    FLD(ev, ORC_F_EXT_B)[i] = rB;
    FLD(ev, ORC_F_EXT_SS)[i] = rS;
    FLD(ev, ORC_F_EXT_Q)[i] = rQ;
    FLD(ev, ORC_F_BULK)[i] = bulk;          /* the reference writes the extensive value here (318) */
    FLD(ev, ORC_F_EZ)[i] = ez;
    for (int d = 0; d < D; ++d) {
      FLD(ev, ORC_F_X + d)[i] = ev->c_r[d][i] + dt*k4col(ev, K4_R + d)[i];
      FLD(ev, ORC_F_UX + d)[i] = ev->c_u[d][i] + dt*k4col(ev, K4_U + d)[i];
    }
    for (int k = 0; k < 5; ++k) FLD(ev, ORC_F_EXT_SHV0 + k)[i] = ev->c_shv[k][i] + dt*k4col(ev, K4_SHV + k)[i];
    for (int k = 0; k < 9; ++k) FLD(ev, ORC_F_EXT_Q0 + k)[i] = ev->c_q[k][i] + dt*k4col(ev, K4_DQ + k)[i];
    for (int d = D; d < 3; ++d) { FLD(ev, ORC_F_X + d)[i] = 0.0; FLD(ev, ORC_F_UX + d)[i] = 0.0; }
  }
  return fail ? ORC_ERR_NEG_ENTROPY : ORC_OK;
}

static int orc_advance_timestep_rk4(orc_event* ev, double dt)
{
  if (!ev->k4) ev->k4 = (double*)calloc((size_t)K4_N * ((size_t)ev->n + 1), sizeof(double));
  else memset(ev->k4, 0, (size_t)K4_N * ((size_t)ev->n + 1) * sizeof(double));
  PHASE(11, orc_cache_state(ev));
  const double t0 = ev->t;
  int st;
  if ((st = orc_get_time_derivatives(ev))) return st;                 /* step 1 */
  PHASE(11, st = orc_rk_update(ev, 0.5*dt, t0));
  if (st) return st;
  if ((st = orc_get_time_derivatives(ev))) return st;                 /* step 2 */
  PHASE(11, st = orc_rk_update(ev, 0.5*dt, t0));
  if (st) return st;
  rk4_accumulate(ev, 1.0/6.0);
  if ((st = orc_get_time_derivatives(ev))) return st;                 /* step 3 */
  PHASE(11, st = orc_rk_update(ev, 0.5*dt, t0));
  if (st) return st;
  rk4_accumulate(ev, 2.0/6.0);
  if ((st = orc_get_time_derivatives(ev))) return st;                 /* step 4 */
  PHASE(11, st = orc_rk_update(ev, dt, t0));
  if (st) return st;
  rk4_accumulate(ev, 2.0/6.0);
  if ((st = orc_get_time_derivatives(ev))) return st;                 /* the last functional call */
  rk4_accumulate(ev, 2.0/6.0);
  PHASE(11, st = rk4_update(ev, dt));
  if (st) return st;
  if (ev->p.particlization) PHASE(12, orc_freeze_out_particles(ev));
  return ORC_OK;
}

int orc_advance_timestep(orc_event* ev, double dt, int rk_order)
{
  if (rk_order == 4) return orc_advance_timestep_rk4(ev, dt);
  if (rk_order != 2) return ORC_ERR_INVALID;
  PHASE(11, orc_cache_state(ev));
  const double t0 = ev->t;
  int st = orc_get_time_derivatives(ev);
  if (st) return st;
  PHASE(11, st = orc_rk_update(ev, 0.5*dt, t0));
  if (st) return st;
  st = orc_get_time_derivatives(ev);
  if (st) return st;
  PHASE(11, st = orc_rk_update(ev, dt, t0));
  if (st) return st;
  if (ev->p.particlization) PHASE(12, orc_freeze_out_particles(ev));
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* BSQHydro::initialize_hydrodynamics (src/BSQHydro.cpp:568-646) after the
 * caller uploaded r, u, input e (fm^-4), input charges, pi^{ij}, Pi. */
int orc_initialize(orc_event* ev)
{
  const int D = ev->p.dim;
  const int n = ev->n;
  const double t0 = ev->p.t0, t2 = t0*t0;
  ev->t = t0;
  /* process_initial_conditions (src/sph_workstation.cpp:1228-1340) */
  for (int i = 0; i < n; ++i) {
    double u[3] = {0,0,0}, uc[3];
    for (int d = 0; d < D; ++d) u[d] = uc[d] = FLD(ev, ORC_F_UX + d)[i];
    if (D == 2) { uc[0] = -uc[0]; uc[1] = -uc[1]; } else { for (int d = 0; d < D-1; ++d) uc[d] = -uc[d]; uc[D-1] *= -t2; }
    double c = 0.0; for (int d = 0; d < D; ++d) c += uc[d]*u[d];
    FLD(ev, ORC_F_GAMMA)[i] = sqrt(1 - c);
    FLD(ev, ORC_F_EXT_S)[i] = FLD(ev, ORC_F_EXT_B)[i] = FLD(ev, ORC_F_EXT_SS)[i] = FLD(ev, ORC_F_EXT_Q)[i] = 1.0;
    FLD(ev, ORC_F_MASS_S)[i] = FLD(ev, ORC_F_MASS_B)[i] = FLD(ev, ORC_F_MASS_SS)[i] = FLD(ev, ORC_F_MASS_Q)[i] = 1.0;
    FLD(ev, ORC_F_T)[i] = 580.0/ORC_HBARC_MEVFM;
    FLD(ev, ORC_F_MUB)[i] = FLD(ev, ORC_F_MUS)[i] = FLD(ev, ORC_F_MUQ)[i] = 0.0;
    ev->I[ORC_I_EOS_ID][i] = 0;
    ev->I[ORC_I_FREEZE][i] = 0;
    ev->I[ORC_I_BTRACK][i] = 0;
  }
  orc_build_neighbors(ev, 0);                           /* initialize_linklist */
  /* initialize_entropy_and_charge_densities (252-375) */
  int st = orc_eos_invert(ev, 1);
  if (st) return st;
  for (int i = 0; i < n; ++i) {
    if (FLD(ev, ORC_F_IN_E)[i] > ev->p.e_freeze && ev->I[ORC_I_FREEZE][i] != 4) ev->I[ORC_I_FREEZE][i] = 0;
    else { ev->I[ORC_I_FREEZE][i] = 4; ev->number_part_fo++; }
  }
  orc_initial_sigma(ev);
  for (int i = 0; i < n; ++i) {
    const double sig = FLD(ev, ORC_F_SIGMA)[i];
    FLD(ev, ORC_F_MASS_S)[i] = FLD(ev, ORC_F_MASS_B)[i] = FLD(ev, ORC_F_MASS_SS)[i] = FLD(ev, ORC_F_MASS_Q)[i] = 1;
    FLD(ev, ORC_F_EXT_S)[i] = FLD(ev, ORC_F_IN_S)[i]/sig;
    FLD(ev, ORC_F_EXT_B)[i] = FLD(ev, ORC_F_IN_B)[i]/sig;
    FLD(ev, ORC_F_EXT_SS)[i] = FLD(ev, ORC_F_IN_SS)[i]/sig;
    FLD(ev, ORC_F_EXT_Q)[i] = FLD(ev, ORC_F_IN_Q)[i]/sig;
    if (FLD(ev, ORC_F_IN_S)[i] < 0.0) ev->I[ORC_I_FREEZE][i] = 4;
  }
  ev->frzc = 0;                                          /* setup_freeze_out */
  ev->fo_step = 0; ev->fo_count = 0; ev->taup = ev->taupp = 0.0;
  for (int k = 0; k < 6; ++k) if (ev->snap[k]) snap_zero(ev->snap[k], n);
  orc_gamma_velocity(ev);                               /* calculate_gamma_and_velocities */
  /* initial_smoothing (394-457) */
  orc_gamma_velocity(ev);
  orc_smooth_fields(ev);
  for (int i = 0; i < n; ++i) {                         /* calculate_extensive_shv (1392-1432) */
    const double sig = FLD(ev, ORC_F_SIGMA)[i];
    int k = 0;
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b, ++k) {
      const double shv = FLD(ev, ORC_F_SHV0 + (a+1)*4 + (b+1))[i];
      FLD(ev, ORC_F_EXT_SHV0 + k)[i] = ev->p.input_initial_shear ? shv/sig : 0.0;
    }
    if (!ev->p.input_initial_shear) for (int q = 0; q < 16; ++q) FLD(ev, ORC_F_SHV0 + q)[i] = 0.0;
    /* set_bulk_Pi (1349-1384) */
    if (ev->p.bulk_from_trace) {
      const double sl = FLD(ev, ORC_F_SIGMA_LAB)[i], g = FLD(ev, ORC_F_GAMMA)[i];
      FLD(ev, ORC_F_EXT_BULK)[i] = (FLD(ev, ORC_F_BULK)[i] - FLD(ev, ORC_F_P)[i])/(sl/g/t0);
    } else { FLD(ev, ORC_F_EXT_BULK)[i] = 0.0; FLD(ev, ORC_F_BULK)[i] = 0.0; }
  }
  if (ev->p.using_diffusion) orc_set_diffusion(ev);     /* set_diffusion (initial_smoothing, 418-419) */
  orc_reset_pi_tensor(ev);
  st = orc_eos_invert(ev, 0);
  if (st) return st;
  /* FreezeOut::check_freeze_out_status with frz1 = 0 */
  check_status(ev);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* SystemState::conservation_* (src/system_state.cpp:738-885) */
double orc_reduce(orc_event* ev, int which)
{
  double acc = 0.0;
  const double t = ev->t;
  for (int i = 0; i < ev->n; ++i) {
    switch (which) {
      case 0: acc += FLD(ev, ORC_F_EXT_S)[i]*FLD(ev, ORC_F_MASS_S)[i]; break;
      case 1: acc += FLD(ev, ORC_F_EXT_B)[i]*FLD(ev, ORC_F_MASS_B)[i]; break;
      case 2: acc += FLD(ev, ORC_F_EXT_SS)[i]*FLD(ev, ORC_F_MASS_SS)[i]; break;
      case 3: acc += FLD(ev, ORC_F_EXT_Q)[i]*FLD(ev, ORC_F_MASS_Q)[i]; break;
      case 4: {
        const double w = FLD(ev, ORC_F_W)[i], p = FLD(ev, ORC_F_P)[i], bulk = FLD(ev, ORC_F_BULK)[i];
        const double g = FLD(ev, ORC_F_GAMMA)[i];
        const double C = w + bulk;
        acc += (C*g*g - p - bulk + FLD(ev, ORC_F_SHV0)[i])*FLD(ev, ORC_F_MASS_S)[i]*t/FLD(ev, ORC_F_SIGMA_LAB)[i];
        break;
      }
      case 5: acc += FLD(ev, ORC_F_EZ)[i]; break;
      case 6: acc += (ev->I[ORC_I_FREEZE][i] >= 4) ? 1.0 : 0.0; break;
      default: break;
    }
  }
  return acc;
}
