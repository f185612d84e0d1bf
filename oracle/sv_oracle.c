/*
 * sv_oracle.c -- CPU oracle (FP64) for mfcm's explicit 2-D Saint-Venant step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see sv_oracle.h for
 * who may load it and what pins each function).  Built with
 * -ffp-contract=off so every product and sum is rounded where the reference
 * rounds it.
 *
 * Reference citations are file:line relative to the reference's src/.
 */
#include "sv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define G 9.81

/* ---- SYCL 1.2.1 built-ins as the reference uses them ------------------ */

/* sycl::sign: 1 for x > 0, -1 for x < 0, the signed zero for +-0, 0 for NaN. */
static inline double sy_sign(double x) {
  if (x > 0.0) return 1.0;
  if (x < 0.0) return -1.0;
  if (x != x) return 0.0;
  return x;
}
/* sycl::min on floating point: y if y < x else x (NaN behaviour unpinned). */
static inline double sy_min(double x, double y) { return (y < x) ? y : x; }
/* sycl::fmax: IEEE maxNum (the non-NaN operand wins), as C fmax. */
static inline double sy_fmax(double x, double y) { return fmax(x, y); }
/* sycl::smoothstep: t = clamp((x - e0) / (e1 - e0), 0, 1); t*t*(3 - 2t). */
static inline double sy_smoothstep(double e0, double e1, double x) {
  double t = (x - e0) / (e1 - e0);
  t = fmin(fmax(t, 0.0), 1.0);
  return t * t * (3.0 - 2.0 * t);
}
/* x^(4/3) for the Manning term (sycl::pow(inv_h, 4/3) at
 * ManningRoughnessSourceTerm.cpp:66-67), evaluated with IEEE operations
 * only so that this oracle and the HIP kernels produce the same bits: five
 * Newton steps for r = x^(-1/3) from an exponent-bit estimate, then
 * (x r)^2.  Within 7 ulp of x^(4/3) on the Manning range (1e-3, 16) (OpenCL
 * allows 16 ulp for pow); the last bits of sycl::pow are parity unpinned. */
double or_pow43(double x) {
  uint64_t i;
  memcpy(&i, &x, 8);
  i = 0x5540000000000000ULL - i / 3u;
  double r;
  memcpy(&r, &i, 8);
  const double third = 1.0 / 3.0;
  for (int k = 0; k < 5; ++k) {
    double r3 = r * r * r;
    double t = 4.0 - x * r3;
    r = r * t * third;
  }
  double c = x * r;
  return c * c;
}

/* sycl::mix: a + (b - a) * t. */
static inline double sy_mix(double a, double b, double t) {
  return a + (b - a) * t;
}

/* ---- Minmod3 (SpatialDerivative/Minmod3.hpp:34-82) ---------------------- */
double or_minmod3(double Ul, double dxl, double Uc, double dxr, double Ur) {
  const double theta = 2.0;                               /* Minmod3.hpp:27 */
  double lhs = NAN;
  if (dxl > 0.01 * dxr) lhs = theta * (Uc - Ul) / dxl;    /* :42-45 */
  double rhs = NAN;
  if (dxr > 0.01 * dxl) rhs = theta * (Ur - Uc) / dxr;    /* :46-49 */
  double central = 0.0;                                   /* :52-56 */
  double total = dxl + dxr;
  if (total > 1e-6) central = (Ur - Ul) / (dxl + dxr);
  double outer;                                           /* :60-70 */
  if (isnan(lhs)) {
    outer = rhs;
  } else if (isnan(rhs)) {
    outer = lhs;
  } else {
    outer = 0.5 * (sy_sign(lhs) + sy_sign(rhs)) * sy_min(fabs(lhs), fabs(rhs));
  }
  if (isnan(outer)) return central;                       /* :74-81 */
  return 0.5 * (sy_sign(outer) + sy_sign(central)) *
         sy_min(fabs(outer), fabs(central));
}

/* ---- Slope passes (SpatialDerivative.hpp:104-122 with the neighbour rules
 * of Mesh/Cartesian2DMesh.hpp:477-520).  X uses (west, east).  Y uses
 * (north, south) with north = row + 1 as the "left" neighbour, so it
 * approximates -d/dy; E7 swaps the pair.  A missing neighbour is the cell
 * itself at distance 0. ------------------------------------------------- */
void or_slope_x(int64_t nx, int64_t ny, double dx, const double* f, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < ny; ++y) {
    for (int64_t x = 0; x < nx; ++x) {
      int64_t i = y * nx + x;
      int64_t il = (x > 0) ? i - 1 : i;
      double dxl = (x > 0) ? dx : 0.0;
      int64_t ir = (x < nx - 1) ? i + 1 : i;
      double dxr = (x < nx - 1) ? dx : 0.0;
      out[i] = or_minmod3(f[il], dxl, f[i], dxr, f[ir]);
    }
  }
}

void or_slope_y(int64_t nx, int64_t ny, double dy, const double* f, double* out,
                int flags) {
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < ny; ++y) {
    for (int64_t x = 0; x < nx; ++x) {
      int64_t i = y * nx + x;
      int64_t in = (y < ny - 1) ? i + nx : i;
      double dyn = (y < ny - 1) ? dy : 0.0;
      int64_t is = (y > 0) ? i - nx : i;
      double dys = (y > 0) ? dy : 0.0;
      if (flags & OR_E7_Y_ORIENT)
        out[i] = or_minmod3(f[is], dys, f[i], dyn, f[in]);
      else
        out[i] = or_minmod3(f[in], dyn, f[i], dys, f[is]);
    }
  }
}

/* ---- One face (SaintVenant/FluxKernel.cpp:66-295) ----------------------- */
void or_face(const or_cellv* Lc, const or_cellv* Rc, int edge, int dir,
             double dx, double out[5]) {
  double xdir = 1.0 - (double)dir;                          /* :76-77 */
  double ydir = (double)dir;
  double zb_L = Lc->zb;                                     /* :80-81 */
  double zb_R = Rc->zb;
  const or_cellv* L = Lc;
  const or_cellv* R = Rc;
  if (zb_L != zb_L) {                                       /* :85-95 */
    L = R;
    if (edge == 1) goto zero;
    edge = -1;
  }
  if (zb_R != zb_R) {                                       /* :96-107 */
    if (edge == -1) goto zero;
    R = L;
    edge = 1;
  }
  {
    double mL = (edge < 0) ? 0.0 : 1.0, mR = (edge > 0) ? 0.0 : 1.0;
    double mLx = (edge < 0 && xdir == 1) ? 0.0 : 1.0;
    double mRx = (edge > 0 && xdir == 1) ? 0.0 : 1.0;
    double mLy = (edge < 0 && ydir == 1) ? 0.0 : 1.0;
    double mRy = (edge > 0 && ydir == 1) ? 0.0 : 1.0;
    double h_L = L->h * mL, h_R = R->h * mR;                /* :111-112 */
    double u_L = L->u * mLx, u_R = R->u * mRx;              /* :116-117 */
    double v_L = L->v * mLy, v_R = R->v * mRy;              /* :121-122 */
    double dhdx_L = L->dhdx * mL, dhdx_R = R->dhdx * mR;    /* :125-128 */
    double dhdy_L = L->dhdy * mL, dhdy_R = R->dhdy * mR;
    double dudx_L = L->dudx * mLx, dudx_R = R->dudx * mRx;  /* :132-135 */
    double dudy_L = L->dudy * mLx, dudy_R = R->dudy * mRx;
    double dvdx_L = L->dvdx * mLy, dvdx_R = R->dvdx * mRy;  /* :139-142 */
    double dvdy_L = L->dvdy * mLy, dvdy_R = R->dvdy * mRy;
    double dzbdx_L = L->dzbdx * mL, dzbdx_R = R->dzbdx * mR; /* :145-148 */
    double dzbdy_L = L->dzbdy * mL, dzbdy_R = R->dzbdy * mR;
    if (edge < 0) zb_L = zb_R + h_R * 2.0;                  /* :152-157 */
    if (edge > 0) zb_R = zb_L + h_L * 2.0;

    /* :161-187 half-cell projections, evaluated left to right */
    double zb_m = zb_L + 0.5 * dx * dzbdx_L * xdir + 0.5 * dx * dzbdy_L * ydir;
    double h_m = h_L + 0.5 * dx * dhdx_L * xdir + 0.5 * dx * dhdy_L * ydir;
    double u_m = u_L + 0.5 * dx * dudx_L * xdir + 0.5 * dx * dudy_L * ydir;
    double v_m = v_L + 0.5 * dx * dvdx_L * xdir + 0.5 * dx * dvdy_L * ydir;
    double zb_p = zb_R + -0.5 * dx * dzbdx_R * xdir + -0.5 * dx * dzbdy_R * ydir;
    double h_p = h_R + -0.5 * dx * dhdx_R * xdir + -0.5 * dx * dhdy_R * ydir;
    double u_p = u_R + -0.5 * dx * dudx_R * xdir + -0.5 * dx * dudy_R * ydir;
    double v_p = v_R + -0.5 * dx * dvdx_R * xdir + -0.5 * dx * dvdy_R * ydir;

    double zb_f = sy_fmax(zb_m, zb_p);                      /* :191 */
    double y_m = zb_m + h_m;                                /* :194-195 */
    double y_p = zb_p + h_p;
    h_m = sy_fmax(h_m, 0.0);                                /* :199-200 */
    h_p = sy_fmax(h_p, 0.0);
    double c_m = sqrt(G * h_m);                             /* :203-204 */
    double c_p = sqrt(G * h_p);

    if (y_m > zb_f && y_p > zb_f) {                         /* :208-234 */
      double spd_m = u_m * xdir + v_m * ydir;
      double spd_p = u_p * xdir + v_p * ydir;
      double hf_m = h_m * spd_m;
      double hf_p = h_p * spd_p;
      double uf_m = u_m * ((1.0 - 0.5 * xdir) * spd_m) + G * h_m * xdir;
      double uf_p = u_p * ((1.0 - 0.5 * xdir) * spd_p) - G * h_p * xdir;
      double vf_m = v_m * ((1.0 - 0.5 * ydir) * spd_m) + G * h_m * ydir;
      double vf_p = v_p * ((1.0 - 0.5 * ydir) * spd_p) - G * h_p * ydir;
      double a = sy_fmax(fabs(spd_p + sy_sign(spd_p) * c_p),
                         fabs(spd_m + sy_sign(spd_m) * c_m));
      out[0] = 0.5 * (hf_p + hf_m) - 0.5 * a * (h_p - h_m);
      out[1] = 0.5 * (uf_p + uf_m) - 0.5 * a * (u_p - u_m);
      out[2] = 0.5 * (vf_p + vf_m) - 0.5 * a * (v_p - v_m);
      out[3] = (zb_m - zb_p) * G;
      out[4] = OR_BRANCH_WET;
    } else if (y_m <= zb_f && y_p <= zb_f) {                /* :235-254 */
      out[0] = 0.0;
      if (zb_p > zb_m) {
        double uf_m = G * h_m * xdir;
        double vf_m = G * h_m * ydir;
        out[1] = 0.5 * uf_m;
        out[2] = 0.5 * vf_m;
        out[3] = -h_m * 0.5 * G;
        out[4] = OR_BRANCH_DRY_R;
      } else {
        double uf_p = G * h_p * xdir;
        double vf_p = G * h_p * ydir;
        out[1] = -0.5 * uf_p;
        out[2] = -0.5 * vf_p;
        out[3] = h_p * 0.5 * G;
        out[4] = OR_BRANCH_DRY_L;
      }
    } else if (y_m > zb_f) {                                /* :255-271 */
      double spd_m = u_m * xdir + v_m * ydir;
      double hf_m = h_m * spd_m;
      double uf_m = u_m * ((1.0 - 0.5 * xdir) * spd_m) + G * h_m * xdir;
      double vf_m = v_m * ((1.0 - 0.5 * ydir) * spd_m) + G * h_m * ydir;
      double a = fabs(spd_m + sy_sign(spd_m) * c_m);
      out[0] = 0.5 * hf_m - 0.5 * a * (-h_m);
      out[1] = 0.5 * uf_m - 0.5 * a * (-u_m);
      out[2] = 0.5 * vf_m - 0.5 * a * (-v_m);
      out[3] = h_p / dx;
      out[4] = OR_BRANCH_LEFT;
    } else {                                                /* :272-289 */
      double spd_p = u_p * xdir + v_p * ydir;
      double hf_p = h_p * spd_p;
      double uf_p = u_p * ((1.0 - 0.5 * xdir) * spd_p) + G * h_p * xdir;
      double vf_p = v_p * ((1.0 - 0.5 * ydir) * spd_p) + G * h_p * ydir;
      double a = fabs(spd_p + sy_sign(spd_p) * c_p);
      out[0] = 0.5 * hf_p - 0.5 * a * (h_p);
      out[1] = 0.5 * uf_p - 0.5 * a * (u_p);
      out[2] = 0.5 * vf_p - 0.5 * a * (v_p);
      out[3] = h_m / dx;
      out[4] = OR_BRANCH_RIGHT;
    }
    return;
  }
zero:                                                       /* :47-60 */
  out[0] = out[1] = out[2] = out[3] = 0.0;
  out[4] = 0.0;
}

/* exp(x) for x <= 0 from IEEE operations only (the HIP kernels evaluate the
 * same expression, so both sides produce the same bits): Cody-Waite
 * reduction by ln 2, degree-12 Taylor polynomial in Horner form, exact
 * scaling by 2^k.  Within 2.3 ulp (checked against 40-digit arithmetic on
 * 25 000 points of [-50, 0]). */
double or_exp_neg(double x) {
  if (x < -708.0) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double log2e = 1.44269504088896338700e+00;
  double k = floor(x * log2e + 0.5);
  double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

/* Storm factor along one axis at coordinate a, centre c: exp(-(a-c)^2/(2 s^2)). */
static double storm_axis(const or_storm* s, double a, double c) {
  double d = a - c;
  return or_exp_neg(-(d * d) / (2.0 * s->sigma * s->sigma));
}

/* Extension E4: peak * exp(-dx^2/(2 sigma^2)) * exp(-dy^2/(2 sigma^2)), the
 * Gaussian written as the product of its two separable factors. */
double or_storm_rate(const or_storm* s, double x, double y, double t) {
  double cx = s->cx0 + s->vx * t;
  double cy = s->cy0 + s->vy * t;
  return s->peak * storm_axis(s, x, cx) * storm_axis(s, y, cy);
}

/* ---- Solver state --------------------------------------------------------- */

typedef struct or_series {
  int n;
  double* t;
  double* v;
  uint8_t* sel; /* NULL = global selection */
} or_series;

#define OR_MAX_STAGES 8
#define OR_MAX_BLOCKS 16

struct or_sim {
  int64_t nx, ny, nc, nf;
  double dx, dy;
  int flags;
  double cn_target;
  double h_dry;                             /* E8 dry depth */
  double u_max;                             /* E9 velocity cap */
  int ns;                                   /* RK stages */
  double a[OR_MAX_STAGES + 1][OR_MAX_STAGES + 1];
  double *zb, *dzbdx, *dzbdy;               /* Constants.cpp:27-44 */
  double *Uh[OR_MAX_STAGES + 1], *Uu[OR_MAX_STAGES + 1], *Uv[OR_MAX_STAGES + 1];
  double *Kh[OR_MAX_STAGES + 1], *Ku[OR_MAX_STAGES + 1], *Kv[OR_MAX_STAGES + 1];
  double *dUdx[3], *dUdy[3];
  double *fh, *fu, *fv, *fz, *fb;           /* Fluxes.hpp: h, u, v, z, branch */
  int manning;
  double *n_sh, *n_dp, *d_sh, *d_dp, *nh, *Sf;
  int ndis, nhead;
  or_series dis[OR_MAX_BLOCKS], head[OR_MAX_BLOCKS];
  double *q0, *q1, *hb0, *hb1;
  int storm;
  or_storm st;
  int eloss;                                /* energy loss: per-cell fx, fy */
  double *fdx, *fdy;
  int infil;                                /* infiltration: rate, capacity (state) */
  double *irate, *icap;
  int eddy;                                 /* 1 constant mu field, 2 Smagorinsky */
  double mu_const, mu_smag;
  double *mu, *d2[4];                       /* d2u/dx2, d2u/dy2, d2v/dx2, d2v/dy2 */
  int nstage;
  or_series stage[OR_MAX_BLOCKS];
  double *sb0, *sb1;                        /* stage (water level) endpoints */
  double ox, oy;
  double start, end, step_dur, dt0, dt;
  int64_t n_acc, n_rej;
  double* trace;
  int64_t trace_cap, trace_len;
};

static double* zalloc(int64_t n) { return (double*)calloc((size_t)n, sizeof(double)); }

or_sim* or_sim_create(int64_t nx, int64_t ny, double dx, double dy, int nstages,
                      const double* tableau, int flags, double cn_target) {
  if (nx < 1 || ny < 1 || nstages < 1 || nstages > OR_MAX_STAGES) return NULL;
  or_sim* S = (or_sim*)calloc(1, sizeof(or_sim));
  S->nx = nx; S->ny = ny; S->nc = nx * ny;
  S->nf = (nx + 1) * ny + (ny + 1) * nx;        /* Cartesian2DMesh.hpp:259-263 */
  S->dx = dx; S->dy = dy; S->flags = flags; S->cn_target = cn_target;
  S->ns = nstages;
  for (int i = 0; i <= nstages; ++i)
    for (int j = 0; j <= nstages; ++j) S->a[i][j] = tableau[i * (nstages + 1) + j];
  S->zb = zalloc(S->nc); S->dzbdx = zalloc(S->nc); S->dzbdy = zalloc(S->nc);
  for (int i = 0; i <= nstages; ++i) {           /* Solver.cpp:45-51 */
    S->Uh[i] = zalloc(S->nc); S->Uu[i] = zalloc(S->nc); S->Uv[i] = zalloc(S->nc);
    S->Kh[i] = zalloc(S->nc); S->Ku[i] = zalloc(S->nc); S->Kv[i] = zalloc(S->nc);
  }
  for (int k = 0; k < 3; ++k) { S->dUdx[k] = zalloc(S->nc); S->dUdy[k] = zalloc(S->nc); }
  S->fh = zalloc(S->nf); S->fu = zalloc(S->nf); S->fv = zalloc(S->nf);
  S->fz = zalloc(S->nf); S->fb = zalloc(S->nf);
  S->q0 = zalloc(S->nc); S->q1 = zalloc(S->nc);
  S->hb0 = zalloc(S->nc); S->hb1 = zalloc(S->nc);
  for (int64_t i = 0; i < S->nc; ++i) { S->hb0[i] = NAN; S->hb1[i] = NAN; }
  S->nh = zalloc(S->nc); S->Sf = zalloc(S->nc);
  S->sb0 = zalloc(S->nc); S->sb1 = zalloc(S->nc);
  for (int64_t i = 0; i < S->nc; ++i) { S->sb0[i] = NAN; S->sb1[i] = NAN; }
  S->start = 0.0; S->end = 0.0; S->step_dur = 60.0; S->dt0 = 1.0; S->dt = 1.0;
  return S;
}

static void free_series(or_series* s) { free(s->t); free(s->v); free(s->sel); }

void or_sim_destroy(or_sim* S) {
  if (!S) return;
  free(S->zb); free(S->dzbdx); free(S->dzbdy);
  for (int i = 0; i <= S->ns; ++i) {
    free(S->Uh[i]); free(S->Uu[i]); free(S->Uv[i]);
    free(S->Kh[i]); free(S->Ku[i]); free(S->Kv[i]);
  }
  for (int k = 0; k < 3; ++k) { free(S->dUdx[k]); free(S->dUdy[k]); }
  free(S->fh); free(S->fu); free(S->fv); free(S->fz); free(S->fb);
  free(S->q0); free(S->q1); free(S->hb0); free(S->hb1);
  free(S->n_sh); free(S->n_dp); free(S->d_sh); free(S->d_dp);
  free(S->nh); free(S->Sf);
  for (int b = 0; b < S->ndis; ++b) free_series(&S->dis[b]);
  for (int b = 0; b < S->nhead; ++b) free_series(&S->head[b]);
  for (int b = 0; b < S->nstage; ++b) free_series(&S->stage[b]);
  free(S->fdx); free(S->fdy); free(S->irate); free(S->icap); free(S->mu);
  for (int k = 0; k < 4; ++k) free(S->d2[k]);
  free(S->sb0); free(S->sb1);
  free(S);
}

/* Constants.cpp:27-44: bed and its two Minmod3 slopes, computed once. */
void or_sim_set_bed(or_sim* S, const double* zb) {
  memcpy(S->zb, zb, sizeof(double) * S->nc);
  or_slope_x(S->nx, S->ny, S->dx, S->zb, S->dzbdx);
  or_slope_y(S->nx, S->ny, S->dy, S->zb, S->dzbdy, S->flags);
}

void or_sim_set_state(or_sim* S, const double* h, const double* u, const double* v) {
  memcpy(S->Uh[0], h, sizeof(double) * S->nc);
  memcpy(S->Uu[0], u, sizeof(double) * S->nc);
  memcpy(S->Uv[0], v, sizeof(double) * S->nc);
}

void or_sim_get_state(const or_sim* S, int idx, double* h, double* u, double* v) {
  if (h) memcpy(h, S->Uh[idx], sizeof(double) * S->nc);
  if (u) memcpy(u, S->Uu[idx], sizeof(double) * S->nc);
  if (v) memcpy(v, S->Uv[idx], sizeof(double) * S->nc);
}

static double* dupd(const double* p, int64_t n) {
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(q, p, sizeof(double) * (size_t)n);
  return q;
}

void or_sim_set_manning(or_sim* S, const double* n_sh, const double* n_dp,
                        const double* d_sh, const double* d_dp) {
  S->manning = 1;
  S->n_sh = dupd(n_sh, S->nc); S->n_dp = dupd(n_dp, S->nc);
  S->d_sh = dupd(d_sh, S->nc); S->d_dp = dupd(d_dp, S->nc);
}

static int add_series(or_sim* S, or_series* arr, int* cnt, int n, const double* t,
                      const double* v, const uint8_t* sel) {
  if (*cnt >= OR_MAX_BLOCKS || n < 2) return -1;
  or_series* s = &arr[(*cnt)++];
  s->n = n;
  s->t = dupd(t, n);
  s->v = dupd(v, n);
  s->sel = NULL;
  if (sel) {
    s->sel = (uint8_t*)malloc((size_t)S->nc);
    memcpy(s->sel, sel, (size_t)S->nc);
  }
  return 0;
}

int or_sim_add_discharge(or_sim* S, int n, const double* t, const double* v,
                         const uint8_t* sel) {
  return add_series(S, S->dis, &S->ndis, n, t, v, sel);
}

int or_sim_add_head(or_sim* S, int n, const double* t, const double* v,
                    const uint8_t* sel) {
  return add_series(S, S->head, &S->nhead, n, t, v, sel);
}

int or_sim_add_stage(or_sim* S, int n, const double* t, const double* v, const uint8_t* sel) {
  return add_series(S, S->stage, &S->nstage, n, t, v, sel);
}

void or_sim_set_energy_loss(or_sim* S, const double* fx, const double* fy) {
  free(S->fdx); free(S->fdy);
  S->fdx = dupd(fx, S->nc); S->fdy = dupd(fy, S->nc);
  S->eloss = 1;
}

void or_sim_set_infiltration(or_sim* S, const double* rate, const double* capacity) {
  free(S->irate); free(S->icap);
  S->irate = dupd(rate, S->nc); S->icap = dupd(capacity, S->nc);
  S->infil = 1;
}

void or_sim_set_eddy_viscosity(or_sim* S, const double* mu, double mu_const, double mu_smag) {
  free(S->mu);
  S->mu = mu ? dupd(mu, S->nc) : zalloc(S->nc);
  for (int k = 0; k < 4; ++k) { free(S->d2[k]); S->d2[k] = zalloc(S->nc); }
  S->mu_const = mu_const; S->mu_smag = mu_smag;
  S->eddy = mu ? 1 : 2;
}

void or_sim_set_storm(or_sim* S, const or_storm* st, double origin_x, double origin_y) {
  S->storm = 1;
  S->st = *st;
  S->ox = origin_x;
  S->oy = origin_y;
}

void or_sim_set_time(or_sim* S, double start, double end, double step_duration,
                     double initial_dt) {
  S->start = start; S->end = end; S->step_dur = step_duration;
  S->dt0 = initial_dt; S->dt = initial_dt;
}

/* TimeParameters.cpp:40-58 */
int64_t or_sim_num_steps(const or_sim* S) {
  double eps = 1e-6;
  double sim_duration = S->end - S->start;
  double fractional = sim_duration / S->step_dur;
  int64_t n = (int64_t)fractional;
  if ((S->step_dur * n) >= sim_duration - eps && (S->step_dur * n) <= sim_duration + eps)
    return n;
  return n + 1;
}

/* Input/TimeSeries.cpp:27-46.  The reference reads one past the end when
 * time >= the last sample; here that case returns NaN instead. */
static double ts_linterp(const or_series* s, double time) {
  int i1 = 1;
  for (; i1 < s->n; ++i1)
    if (s->t[i1] > time) break;
  if (i1 >= s->n) return NAN;
  int i0 = i1 - 1;
  double t0 = s->t[i0], t1 = s->t[i1];
  double v0 = s->v[i0], v1 = s->v[i1];
  double dvdt = (v1 - v0) / (t1 - t0);
  double dt = time - t0;
  return v0 + dvdt * dt;
}

/* Field.cpp:91-96: operator=(scalar) is (*this) *= 0; (*this) += v. */
static void field_assign(double* f, int64_t n, double v) {
  for (int64_t i = 0; i < n; ++i) { f[i] *= 0.0; f[i] += v; }
}

/* Solver.cpp:119-129 -> DischargeBoundarySourceTerm.hpp:114-148 and
 * HeadBoundarySourceTerm.hpp:114-142: per-step endpoint fields. */
void or_sim_start_new_step(or_sim* S, double time_now) {
  const double sd = S->step_dur;
  if (S->ndis > 0 || S->storm) {
    field_assign(S->q0, S->nc, 0.0);
    field_assign(S->q1, S->nc, 0.0);
    for (int b = 0; b < S->ndis; ++b) {
      double a0 = ts_linterp(&S->dis[b], time_now);
      double a1 = ts_linterp(&S->dis[b], time_now + sd);
      const uint8_t* sel = S->dis[b].sel;
      for (int64_t i = 0; i < S->nc; ++i)
        if (!sel || sel[i]) { S->q0[i] += a0; S->q1[i] += a1; }
    }
    if (S->storm) {
      for (int64_t y = 0; y < S->ny; ++y)
        for (int64_t x = 0; x < S->nx; ++x) {
          int64_t i = y * S->nx + x;
          double cx = S->ox + ((double)x + 0.5) * S->dx;
          double cy = S->oy + ((double)y + 0.5) * S->dy;
          S->q0[i] += or_storm_rate(&S->st, cx, cy, time_now);
          S->q1[i] += or_storm_rate(&S->st, cx, cy, time_now + sd);
        }
    }
  }
  /* StageBoundarySourceTerm.hpp:114-142: water-level endpoints */
  if (S->nstage > 0) {
    field_assign(S->sb0, S->nc, 0.0);
    field_assign(S->sb1, S->nc, 0.0);
    for (int b = 0; b < S->nstage; ++b) {
      double a0 = ts_linterp(&S->stage[b], time_now);
      double a1 = ts_linterp(&S->stage[b], time_now + sd);
      const uint8_t* sel = S->stage[b].sel;
      for (int64_t i = 0; i < S->nc; ++i)
        if (!sel || sel[i]) { S->sb0[i] = a0; S->sb1[i] = a1; }
    }
  }
  if (S->nhead > 0) {
    field_assign(S->hb0, S->nc, 0.0);
    field_assign(S->hb1, S->nc, 0.0);
    for (int b = 0; b < S->nhead; ++b) {
      double a0 = ts_linterp(&S->head[b], time_now);
      double a1 = ts_linterp(&S->head[b], time_now + sd);
      const uint8_t* sel = S->head[b].sel;
      for (int64_t i = 0; i < S->nc; ++i)
        if (!sel || sel[i]) { S->hb0[i] = a0; S->hb1[i] = a1; }
    }
  }
}

static inline void gather_cell(const or_sim* S, const double* h, const double* u,
                               const double* v, int64_t i, or_cellv* c) {
  c->h = h[i]; c->u = u[i]; c->v = v[i]; c->zb = S->zb[i];
  c->dzbdx = S->dzbdx[i]; c->dzbdy = S->dzbdy[i];
  c->dhdx = S->dUdx[0][i]; c->dudx = S->dUdx[1][i]; c->dvdx = S->dUdx[2][i];
  c->dhdy = S->dUdy[0][i]; c->dudy = S->dUdy[1][i]; c->dvdy = S->dUdy[2][i];
}

/* Fluxes.cpp:48-65 with the adjacency of Cartesian2DMesh.hpp:379-435. */
static void face_pass(or_sim* S, const double* h, const double* u, const double* v) {
  const int64_t nx = S->nx, ny = S->ny;
  const int64_t nxf = (nx + 1) * ny;
#pragma omp parallel for schedule(static)
  for (int64_t fid = 0; fid < S->nf; ++fid) {
    int64_t lhs, rhs;
    int edge, dir;
    double d;
    int west_outflow = 0;
    if (fid < nxf) {
      dir = 0; d = S->dx;
      int64_t fx = fid % (nx + 1), fy = fid / (nx + 1);
      if (fx < nx) {
        rhs = fy * nx + fx;
        if (fx > 0) { edge = 0; lhs = fy * nx + (fx - 1); }
        else { edge = -1; lhs = rhs; west_outflow = (S->flags & OR_E2_WEST_OUTFLOW) != 0; }
      } else {
        lhs = fy * nx + (fx - 1); rhs = lhs; edge = 1;
      }
    } else {
      dir = 1; d = S->dy;
      int64_t f2 = fid - nxf;
      int64_t fx = f2 % nx, fy = f2 / nx;
      if (fy < ny) {
        rhs = fy * nx + fx;
        if (fy > 0) { edge = 0; lhs = (fy - 1) * nx + fx; }
        else { edge = -1; lhs = rhs; }
      } else {
        lhs = (fy - 1) * nx + fx; rhs = lhs; edge = 1;
      }
    }
    or_cellv L, R;
    gather_cell(S, h, u, v, lhs, &L);
    gather_cell(S, h, u, v, rhs, &R);
    if (west_outflow) {
      /* E2: zero-gradient ghost that only lets water leave through x = 0. */
      memset(&L, 0, sizeof(L));
      L.h = R.h; L.u = fmin(R.u, 0.0); L.v = R.v; L.zb = R.zb;
      edge = 0;
    }
    double out[5];
    or_face(&L, &R, edge, dir, d, out);
    S->fh[fid] = out[0]; S->fu[fid] = out[1]; S->fv[fid] = out[2];
    S->fz[fid] = out[3]; S->fb[fid] = out[4];
  }
}

/* TemporalDerivativeKernel.cpp:47-99 (cell pass, overwrites dUdt) with the
 * face numbering of Cartesian2DMesh.hpp:447-460. */
static void cell_pass(or_sim* S, const double* h, double* dh, double* du, double* dv) {
  const int64_t nx = S->nx, ny = S->ny;
  const int64_t nxf = (nx + 1) * ny;
  const double dx = S->dx, dy = S->dy;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    int64_t cx = c % nx, cy = c / nx;
    int64_t fw = cy * (nx + 1) + cx, fe = fw + 1;
    int64_t fs = nxf + cy * nx + cx, fn = fs + nx;
    double dhdt = (S->fh[fw] - S->fh[fe]) / dx + (S->fh[fs] - S->fh[fn]) / dy;
    double dudt = (S->fu[fw] - S->fu[fe]) / dx + (S->fu[fs] - S->fu[fn]) / dy;
    double dvdt = (S->fv[fw] - S->fv[fe]) / dx + (S->fv[fs] - S->fv[fn]) / dy;
    dudt += S->dUdx[0][c] * -G;                             /* :66-67 */
    dvdt += S->dUdy[0][c] * -G;
    double dzdx = S->dzbdx[c];                              /* :72-79 */
    if (fabs(dzdx) > h[c] / dx) dzdx = sy_sign(dzdx) * h[c] / dx;
    double dzdy = S->dzbdy[c];
    if (fabs(dzdy) > h[c] / dy) dzdy = sy_sign(dzdy) * h[c] / dy;
    double dudt_bed = -G * dzdx;                            /* :80-81 */
    double dvdt_bed = -G * dzdy;
    dudt += (S->fz[fw] - S->fz[fe]) / dx;                   /* :86-87 */
    dvdt += (S->fz[fs] - S->fz[fn]) / dy;
    dudt += dudt_bed;                                       /* :89-90 */
    dvdt += dvdt_bed;
    dh[c] = dhdt; du[c] = dudt; dv[c] = dvdt;
  }
}

/* ManningRoughnessSourceTerm.cpp:49-91. */
static void manning_pass(or_sim* S, const double* h_, const double* u_, const double* v_,
                         double* du, double* dv, double timestep) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h = h_[c], u = u_[c], v = v_[c];
    double n = sy_mix(S->n_sh[c], S->n_dp[c], sy_smoothstep(S->d_sh[c], S->d_dp[c], h));
    S->nh[c] = n;
    if (h > 1e-6) {
      double inv_h = h / (h * h + 1e-3);
      double Sf = n * n * sqrt(u * u + v * v) * or_pow43(inv_h);
      S->Sf[c] = Sf;
      double dudt = -G * Sf * u;
      double dvdt = -G * Sf * v;
      if (fabs(dudt * timestep) > fabs(u) && sy_sign(dudt * timestep) != sy_sign(u))
        dudt = -u / timestep;
      if (fabs(dvdt * timestep) > fabs(v) && sy_sign(dvdt * timestep) != sy_sign(v))
        dvdt = -v / timestep;
      du[c] += dudt;
      dv[c] += dvdt;
    } else {
      S->Sf[c] = 0.0;
    }
  }
}

/* DischargeBoundarySourceTerm.cpp:43-67 (E1 flips the sign in the guard). */
static void discharge_pass(or_sim* S, const double* h_, double* dh, double timestep,
                           double time_now) {
  const double step_length = S->step_dur;
  const int e1 = (S->flags & OR_E1_RAIN_GUARD) != 0;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double q0 = S->q0[c], q1 = S->q1[c];
    if (q0 != 0.0 || q1 != 0.0) {
      double h = h_[c];
      double dqdt = (q1 - q0) / step_length;
      double qnow = q0 + time_now * dqdt;
      double qnext = qnow + timestep * dqdt;
      double dhdt = 0.5 * (qnow + qnext);
      int dry = e1 ? (h + dhdt * timestep <= 0.0) : (h - dhdt * timestep <= 0.0);
      if (dry) dhdt = -h / timestep;
      dh[c] += dhdt;
    }
  }
}

/* HeadBoundarySourceTerm.cpp:43-66 (overwrites dh where the target is set). */
static void head_pass(or_sim* S, const double* h_, double* dh, double timestep,
                      double time_now) {
  const double step_length = S->step_dur;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h0 = S->hb0[c], h1 = S->hb1[c];
    if (!(h0 != h0) && !(h1 != h1)) {
      double h = h_[c];
      double dhbdt = (h1 - h0) / step_length;
      double now = h0 + time_now * dhbdt;
      double next = now + timestep * dhbdt;
      double target = 0.5 * (now + next);
      if (target <= 0.0) target = 0.0;
      dh[c] = (target - h) / timestep;
    }
  }
}

/* EnergyLossSourceTerm.cpp:59-88: quadratic loss with per-cell fx, fy. */
static void eloss_pass(or_sim* S, const double* h_, const double* u_, const double* v_,
                       double* du, double* dv, double timestep) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h = h_[c], u = u_[c], v = v_[c];
    if (h > 1e-6) {
      double hyp = sqrt(u * u + v * v);
      double dudt = -0.5 * u * S->fdx[c] * hyp;
      double dvdt = -0.5 * v * S->fdy[c] * hyp;
      if (fabs(dudt * timestep) > fabs(u) && sy_sign(dudt * timestep) != sy_sign(u))
        dudt = -u / timestep;
      if (fabs(dvdt * timestep) > fabs(v) && sy_sign(dvdt * timestep) != sy_sign(v))
        dvdt = -v / timestep;
      du[c] += dudt;
      dv[c] += dvdt;
    }
  }
}

/* InfiltrationSourceTerm.cpp:42-64: the capacity is state and shrinks at
 * every application (every stage, rejected timesteps included). */
static void infil_pass(or_sim* S, const double* h_, double* dh, double timestep) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h = h_[c];
    double d = S->irate[c] * timestep;
    if (d > h) d = h;
    if (d > S->icap[c]) d = S->icap[c];
    dh[c] -= d / timestep;
    S->icap[c] -= d;
  }
}

/* EddyViscositySourceTerm.hpp:216-256 and .cpp:44-98: second derivatives
 * are Minmod3 slopes of the first-derivative fields (x of dU/dx, y of
 * dU/dy); Smagorinsky mu from the first derivatives.  pow(a, 2) is a*a. */
static void eddy_pass(or_sim* S, const double* h_, const double* u_, const double* v_,
                      double* du, double* dv, double timestep) {
  or_slope_x(S->nx, S->ny, S->dx, S->dUdx[1], S->d2[0]);
  or_slope_y(S->nx, S->ny, S->dy, S->dUdy[1], S->d2[1], S->flags);
  or_slope_x(S->nx, S->ny, S->dx, S->dUdx[2], S->d2[2]);
  or_slope_y(S->nx, S->ny, S->dy, S->dUdy[2], S->d2[3], S->flags);
  if (S->eddy == 2) {
    const double area = S->dx * S->dy;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < S->nc; ++c) {
      double dudx = S->dUdx[1][c], dudy = S->dUdy[1][c];
      double dvdx = S->dUdx[2][c], dvdy = S->dUdy[2][c];
      double t = fabs(dudy) + fabs(dvdx);
      S->mu[c] = S->mu_const + S->mu_smag * area *
                 sqrt(dudx * dudx + dvdy * dvdy + 0.5 * (t * t));
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h = h_[c], u = u_[c], v = v_[c];
    if (h > 1e-6) {
      double dudt = -S->mu[c] * (S->d2[0][c] + S->d2[1][c]);
      double dvdt = -S->mu[c] * (S->d2[2][c] + S->d2[3][c]);
      if (fabs(dudt * timestep) > fabs(u) && sy_sign(dudt * timestep) != sy_sign(u))
        dudt = -u / timestep;
      if (fabs(dvdt * timestep) > fabs(v) && sy_sign(dvdt * timestep) != sy_sign(v))
        dvdt = -v / timestep;
      du[c] += dudt;
      dv[c] += dvdt;
    }
  }
}

/* StageBoundarySourceTerm.cpp:43-68: relax the depth to water level minus
 * bed (overwrites dh where the level is set). */
static void stage_pass(or_sim* S, const double* h_, double* dh, double timestep,
                       double time_now) {
  const double step_length = S->step_dur;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < S->nc; ++c) {
    double z = S->zb[c];
    double h0 = S->sb0[c] - z, h1 = S->sb1[c] - z;
    if (!(h0 != h0) && !(h1 != h1)) {
      double h = h_[c];
      double dhbdt = (h1 - h0) / step_length;
      double now = h0 + time_now * dhbdt;
      double next = now + timestep * dhbdt;
      double target = 0.5 * (now + next);
      if (target <= 0.0) target = 0.0;
      dh[c] = (target - h) / timestep;
    }
  }
}

/* Solver.cpp:77-114: slopes -> faces -> cell pass -> sources -> boundaries.
 * Sources and boundaries run in config order; the order restated here is
 * Manning, energy loss, infiltration, eddy viscosity, then the discharge,
 * head and stage blocks. */
void or_sim_update_dudt(or_sim* S, int s, double t_sub, double dt) {
  const double* h = S->Uh[s];
  const double* u = S->Uu[s];
  const double* v = S->Uv[s];
  or_slope_x(S->nx, S->ny, S->dx, h, S->dUdx[0]);          /* State.cpp:187-192 */
  or_slope_x(S->nx, S->ny, S->dx, u, S->dUdx[1]);
  or_slope_x(S->nx, S->ny, S->dx, v, S->dUdx[2]);
  or_slope_y(S->nx, S->ny, S->dy, h, S->dUdy[0], S->flags);
  or_slope_y(S->nx, S->ny, S->dy, u, S->dUdy[1], S->flags);
  or_slope_y(S->nx, S->ny, S->dy, v, S->dUdy[2], S->flags);
  face_pass(S, h, u, v);
  cell_pass(S, h, S->Kh[s], S->Ku[s], S->Kv[s]);
  if (S->manning) manning_pass(S, h, u, v, S->Ku[s], S->Kv[s], dt);
  if (S->eloss) eloss_pass(S, h, u, v, S->Ku[s], S->Kv[s], dt);
  if (S->infil) infil_pass(S, h, S->Kh[s], dt);
  if (S->eddy) eddy_pass(S, h, u, v, S->Ku[s], S->Kv[s], dt);
  if (S->ndis > 0 || S->storm) discharge_pass(S, h, S->Kh[s], dt, t_sub);
  if (S->nhead > 0) head_pass(S, h, S->Kh[s], dt, t_sub);
  if (S->nstage > 0) stage_pass(S, h, S->Kh[s], dt, t_sub);
}

/* RungeKutta.hpp:371-405: y_i = ((y_0 + k_0*(a dt)) + k_1*(a dt)) + ... */
void or_sim_update_y(or_sim* S, int step, double dt) {
  if (step <= 0) return;
  double* yh = S->Uh[step];
  double* yu = S->Uu[step];
  double* yv = S->Uv[step];
  memcpy(yh, S->Uh[0], sizeof(double) * S->nc);
  memcpy(yu, S->Uu[0], sizeof(double) * S->nc);
  memcpy(yv, S->Uv[0], sizeof(double) * S->nc);
  for (int col = 0; col < step; ++col) {
    double s = S->a[step][col + 1] * dt;  /* coeffs().a(step+1, col+1) * timestep */
    const double* kh = S->Kh[col];
    const double* ku = S->Ku[col];
    const double* kv = S->Kv[col];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < S->nc; ++c) {
      yh[c] += kh[c] * s;
      yu[c] += ku[c] * s;
      yv[c] += kv[c] * s;
    }
  }
  /* E8: a completed stage state carries no velocity where h <= h_dry. */
  if (S->flags & OR_E8_DRY_VELOCITY) {
    const double hd = S->h_dry;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < S->nc; ++c)
      if (yh[c] <= hd) { yu[c] = 0.0; yv[c] = 0.0; }
  }
  /* E9: each velocity component of a completed stage state within +-u_max. */
  if (S->flags & OR_E9_VELOCITY_CAP) {
    const double um = S->u_max;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < S->nc; ++c) {
      yu[c] = fmin(fmax(yu[c], -um), um);
      yv[c] = fmin(fmax(yv[c], -um), um);
    }
  }
}

void or_sim_set_dry_depth(or_sim* S, double h_dry) { S->h_dry = h_dry; }
void or_sim_set_velocity_cap(or_sim* S, double u_max) { S->u_max = u_max; }

/* State.cpp:199-229: max over cells of dt ((|u| + c)/dx + (|v| + c)/dy). */
double or_sim_control_number(or_sim* S, int s, double timestep) {
  const double* H = S->Uh[s];
  const double* U = S->Uu[s];
  const double* V = S->Uv[s];
  const double dx = S->dx, dy = S->dy;
  double m = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m)
  for (int64_t c = 0; c < S->nc; ++c) {
    double h = sy_fmax(H[c], 0.0);
    double u = fabs(U[c]);
    double v = fabs(V[c]);
    double cc = sqrt(G * h);
    double cn = timestep * (((u + cc) / dx) + ((v + cc) / dy));
    if (cn > m) m = cn;
  }
  return m;
}

static void trace_push(or_sim* S, double t, double dt, double cn, int rep) {
  if (!S->trace || S->trace_len >= S->trace_cap) { S->trace_len++; return; }
  double* p = S->trace + 4 * S->trace_len++;
  p[0] = t; p[1] = dt; p[2] = cn; p[3] = rep;
}

/* RungeKutta.cpp:143-167.  The reference hard-codes cn_target = 1.0 and its
 * formulas dt/(cn*1.1) and dt/cn assume that target; extension E3 (a lower
 * target such as 0.5) applies the same rules to cn / cn_target, which is the
 * identical computation when cn_target == 1.0. */
int or_sim_do_timestep(or_sim* S, double local_time, double dt, double* new_dt) {
  for (int sub = 1; sub <= S->ns; ++sub) {
    double t_sub = local_time + S->a[sub - 1][0] * dt;  /* c(substep) */
    or_sim_update_dudt(S, sub - 1, t_sub, dt);
    or_sim_update_y(S, sub, dt);
  }
  double cn_raw = or_sim_control_number(S, S->ns, dt);
  double cn = (S->cn_target == 1.0) ? cn_raw : cn_raw / S->cn_target;
  const double cn_target = 1.0;
  int repeat;
  if (cn > cn_target) {
    repeat = 1;
    if (cn > 5.0 * cn_target) *new_dt = dt / 5.0;
    else *new_dt = dt / (cn * 1.1);
  } else if (cn > 0.9 * cn_target) {
    repeat = 0; *new_dt = dt;
  } else {
    repeat = 0; *new_dt = dt / cn;
  }
  trace_push(S, local_time, dt, cn_raw, repeat);
  return repeat;
}

/* RungeKutta.hpp:428-431: y_0 <- y_N. */
void or_sim_accept(or_sim* S) {
  memcpy(S->Uh[0], S->Uh[S->ns], sizeof(double) * S->nc);
  memcpy(S->Uu[0], S->Uu[S->ns], sizeof(double) * S->nc);
  memcpy(S->Uv[0], S->Uv[S->ns], sizeof(double) * S->nc);
}

/* TemporalScheme.hpp:112-149. */
void or_sim_step(or_sim* S) {
  double t_local = 0.0;
  double step = S->step_dur;
  while (t_local < step) {
    double new_dt;
    int repeat = or_sim_do_timestep(S, t_local, S->dt, &new_dt);
    if (repeat) {
      S->n_rej++;
    } else {
      S->n_acc++;
      t_local += S->dt;
      or_sim_accept(S);
    }
    S->dt = new_dt;
    if (t_local >= step) return;
    else if (t_local + S->dt > step) S->dt = step - t_local;
    else if (t_local + 2.0 * S->dt > step) S->dt = (step - t_local) * 0.4;
  }
}

/* TemporalScheme.hpp:244-276 over outer steps [first, first + count). */
void or_sim_run_steps(or_sim* S, int64_t first, int64_t count) {
  for (int64_t i = first; i < first + count; ++i) {
    double time_now = S->start + (double)i * S->step_dur;
    or_sim_start_new_step(S, time_now);
    or_sim_step(S);
  }
}

void or_sim_solve(or_sim* S) {
  S->dt = S->dt0;
  or_sim_run_steps(S, 0, or_sim_num_steps(S));
}

double or_sim_dt(const or_sim* S) { return S->dt; }
void or_sim_set_dt(or_sim* S, double dt) { S->dt = dt; }

void or_sim_stats(const or_sim* S, int64_t* stats) {
  stats[0] = S->n_acc;
  stats[1] = S->n_rej;
}

/* Sum of h * cell_area (Output/Measure.hpp:137-153) in FP64 with per-row
 * Neumaier compensation and a fixed row order. */
double or_sim_volume(const or_sim* S, int s) {
  const double* H = S->Uh[s];
  double area = S->dx * S->dy;
  double sum = 0.0, comp = 0.0;
  for (int64_t c = 0; c < S->nc; ++c) {
    double x = H[c] * area;
    double t = sum + x;
    if (fabs(sum) >= fabs(x)) comp += (sum - t) + x;
    else comp += (x - t) + sum;
    sum = t;
  }
  return sum + comp;
}

void or_sim_set_trace(or_sim* S, double* buf, int64_t cap) {
  S->trace = buf; S->trace_cap = cap; S->trace_len = 0;
}
int64_t or_sim_trace_len(const or_sim* S) { return S->trace_len; }

double* or_sim_array(or_sim* S, int which, int idx) {
  switch (which) {
    case 0: return S->Uh[idx];
    case 1: return S->Uu[idx];
    case 2: return S->Uv[idx];
    case 3: return S->Kh[idx];
    case 4: return S->Ku[idx];
    case 5: return S->Kv[idx];
    case 6: case 7: case 8: return S->dUdx[which - 6];
    case 9: case 10: case 11: return S->dUdy[which - 9];
    case 12: return S->zb;
    case 13: return S->dzbdx;
    case 14: return S->dzbdy;
    case 15: return S->fh;
    case 16: return S->fu;
    case 17: return S->fv;
    case 18: return S->fz;
    case 19: return S->fb;
    case 20: return S->q0;
    case 21: return S->q1;
    case 22: return S->hb0;
    case 23: return S->hb1;
    case 24: return S->nh;
    case 25: return S->Sf;
    case 26: return S->icap;
    case 27: return S->mu;
    case 28: case 29: case 30: case 31: return S->d2[which - 28];
    case 32: return S->sb0;
    case 33: return S->sb1;
    default: return NULL;
  }
}
