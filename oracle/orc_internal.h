/* Internal declarations of the CPU oracle (test infrastructure only). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include "ccake_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/constants.h:15-16 */
#define ORC_HBARC_MEVFM 197.3269804
#define ORC_HBARC_GEVFM 0.1973269804
#define ORC_PI 3.1415926535897932384626433
/* include/eos_header.h:16-17, include/rootfinder.h:36-37 */
#define ORC_TINY 1e-25
#define ORC_TBQS_INF 1e10
#define ORC_RF_TOL 1e-10

int orc_nthreads(void);

/* -------- EoS chain members (src/eos_initialization.cpp:216-501) -------- */
enum { ORC_M_TABLE = 0, ORC_M_TANH, ORC_M_CONF, ORC_M_DIAG };

typedef struct orc_member {
  int kind;
  double c, T0, muB0, muS0, muQ0, Tc, Ts;
  double mins[4], maxs[4];            /* tbqs_minima / tbqs_maxima           */
  /* table member */
  int dims[4];
  double gmin[4], gmax[4], gdel[4];
  double* tab;                        /* node-major, 17 fields per node      */
} orc_member;

/* flat pre-inverted table (EoS_Interpolator, include/eos_interpolator.h),
 * variables in the reference's eos_variables order */
enum { PV_T = 0, PV_MUB, PV_MUQ, PV_MUS, PV_E, PV_P, PV_CS2, PV_DWDS, PV_DWDB, PV_DWDQ, PV_DWDS_S, PV_NUM };
typedef struct orc_preinv {
  int Ns, NB;
  double ls0, dls, lb0, dlb;       /* log10 axis start and stride */
  double *s_axis, *nB_axis;        /* fm^-3 */
  double* v[PV_NUM];               /* Ns*NB each, s slowest, fm units */
} orc_preinv;

struct orc_eos {
  int kind;
  int n;
  orc_member m[4];
  orc_preinv* pre;
};

/* BSQ diffusion (orc_diff.c) */
void orc_diffusion_transport(const orc_params* P, const double* th, double kappa[3][3], double tauq[3][3]);
void orc_regulate_diffusion(orc_event* ev);
void orc_set_diffusion(orc_event* ev);
void orc_mrf_diffusion(orc_event* ev, int i, const double M_S_in[3], double F_S, const double R_S[3],
                       double F_q0[3], double R_q0[3][3], double M_q0[3][3],
                       double F_q[3][3], double M_q[3][3][3], double R_q[3][3][3]);

orc_preinv* orc_preinv_create(const orc_eos_desc* d);
void orc_preinv_destroy(orc_preinv* p);
double orc_preinv_efreeze(const orc_preinv* p, double T_fm);
int orc_preinv_invert(orc_event* ev, int mode);

/* chi indices inside the 17-field thermo vector (src/eos_table.cpp:30-34) */
enum { X_P = 0, X_S, X_B, X_SS, X_Q, X_E, X_CS2, X_BB, X_QQ, X_SSS, X_BQ, X_BS,
       X_QS, X_TB, X_TQ, X_TS, X_TT };

void orc_member_full(const orc_member* m, const double x[4], double r[17]);

/* -------- events -------- */
struct orc_event {
  orc_params p;
  orc_eos* eos;
  int n;
  double t;
  double* F[ORC_NUM_FIELDS];
  int32_t* I[ORC_NUM_IFIELDS];
  /* RK cache (src/evolver.cpp:124-168) */
  double* c_u[3];
  double* c_r[3];
  double* c_s, *c_B, *c_S, *c_Q, *c_bulk, *c_E0;
  double* c_shv[5];
  double* c_q[9];        /* extensive diffusion (evolver.cpp:132-152) */
  /* RK4 accumulator k (Evolver::k, src/evolver.cpp:171-212): u, r, s, B, S,
   * Q, bulk, E0, 5 shear, 9 diffusion; allocated at the first RK4 step */
  double* k4;
  /* neighbour list (CSR, j ascending) */
  int64_t* nb_off;
  int32_t* nb_idx;
  int64_t nb_cap;
  int64_t n_pairs;
  /* freeze-out hypersurface (include/freeze_out.h:29-35, 165-222): the six
   * snapshot sets and the appended rows (particlization on) */
  struct orc_snap* snap[6];      /* frz1, frz2, fback, fback2, fback3, fback4 */
  double* fo_rows;               /* n x ORC_FO_ROW */
  int fo_count, fo_step;
  double taup, taupp;
  /* freeze-out bookkeeping (include/freeze_out.h:165-222) */
  int frzc;
  double* fback_e[4];    /* fback, fback2, fback3, fback4 energy snapshots */
  double* fback_gradP0[4];
  double* frz1_gradP0, *frz2_gradP0;
  int number_part_fo;
};

#define FLD(ev, f) ((ev)->F[(f)])

/* One freeze_array for D = 2 (include/freeze_out.h:29-35): the members the
 * hypersurface reads; diffusion is off and stays zero. */
typedef struct orc_snap {
  double* shear[3][3];
  double* r[2]; double* u[2]; double* gradP[2]; double* gradE[2];
  double* s; double* e; double* rhoB; double* rhoS; double* rhoQ;
  double* T; double* muB; double* muS; double* muQ;
  double* theta; double* bulk; double* sigma_lab; double* shear33; double* shv_nabla_u;
  double* wfz; double* cs2fz;
  double* diff[12];      /* q^mu of B, S, Q (freeze_array diff) */
} orc_snap;

#endif
