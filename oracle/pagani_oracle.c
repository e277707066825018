/*
 * pagani_oracle.c -- CPU restatement of the PAGANI adaptive cubature path.
 *
 * TEST INFRASTRUCTURE ONLY (see pagani_oracle.h).  Plain C99 + OpenMP.
 * Region evaluation is parallel over regions (deterministic: every region
 * is independent); every sum is a sequential left-to-right loop.
 *
 * Reference files restated here (paths relative to the reference tree):
 *   cuda/pagani/quad/GPUquad/Rule.cuh            rule tables, orbits
 *   cuda/pagani/quad/GPUquad/Sample.cuh          per-region rule sums
 *   cuda/pagani/quad/GPUquad/Phases.cuh          region setup, RefineError
 *   cuda/pagani/quad/GPUquad/Sub_regions.cuh     uniform split
 *   cuda/pagani/quad/GPUquad/Sub_region_filter.cuh / Sub_region_splitter.cuh
 *   cuda/pagani/quad/GPUquad/heuristic_classifier.cuh
 *   cuda/pagani/quad/GPUquad/PaganiUtils.cuh     finished sums, convergence
 *   cuda/pagani/quad/GPUquad/Workspace.cuh       the iteration loop
 *   cuda/pagani/demos/compute_genz_integrals.cuh closed forms
 */
#include "pagani_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================== */
/* Rule tables                                                               */
/* ======================================================================== */

/* Points per region, quad.h:68-75 / Rule.cuh:409-411. */
int orc_feval(int d)
{
  if (d < 2 || d > ORC_MAXDIM) return -1;
  return 1 + 2 * d + 2 * d + 2 * d + 2 * d + 2 * d * (d - 1) +
         4 * d * (d - 1) + 4 * d * (d - 1) * (d - 2) / 3 + (1 << d);
}

static void set_counts(int d, int* c)
{
  /* Rule.cuh:155-163 */
  c[0] = 1;
  c[1] = c[2] = c[3] = c[4] = 2 * d;
  c[5] = 2 * d * (d - 1);
  c[6] = 4 * d * (d - 1);
  c[7] = 4 * d * (d - 1) * (d - 2) / 3;
  c[8] = 1 << d;
}

/* Degree-9 weights as polynomials in d, Rule.cuh:36-147.  The operation
 * order of every expression is kept so the doubles are the reference's. */
static void rule9_weights(int d, double w[ORC_NSETS][ORC_NRULES])
{
  const double D = (double)d;
  const double a3 = -.002361170967785511788400941242259231309691;
  const double a2 = .1141539002385732526821323741697655347686;
  const double a1 = -.6383392007670238909386026193674701393074;
  const double a0 = .7484998850468520800423030047583803945205;
  const double b3 = -.001432401703339912514196154599769007103671;
  const double b2 = .05747150786448972594860897296200006759892;
  const double b1 = -.1422510457143424323449521620935950679394;
  const double b0 = -.06287502873828697998942424881040490136987;
  const double c1 = .2545911332489590890011611142429070613156;
  const double e2 = -1.207328566678236261002219995185143356737;
  const double e1 = .8956736576416067650809467826488567200939;
  const double f1 = -.3647935698604914666100134551377381205297;
  const double g2 = .003541756451678267682601411863388846964536;
  const double g1 = -.07260936739589367960492815865074633743652;
  const double g0 = .1055749162521899101218622863269817454540;
  const double h2 = .002148602555009868771294231899653510655506;
  const double h1 = -.03226856389295394999786630399875134318006;
  const double h0 = .01063678399023121748083624225818915724455;
  const double k0 = .01468910249614349017540783437728097691502;
  const double m1 = .5113470834646759143109387357149329909126;
  const double m0 = .4597644812080634464633352781605214342691;
  const double n0 = .1823967849302457333050067275688690602649;
  const double p1 = -.04508628929435784075980562738240804429658;
  const double p0 = .2141588352435279340097929526588394300172;
  const double q1 = -.02735154652654564472203690086290223507436;
  const double q0 = .05494106704871123410060080562462135546101;
  const double r0 = .1193759620257077529708962121565290178730;
  const double s1 = .6508951939192025059314756320878023215278;
  const double s0 = .1474493982943446016775696826942585013243;
  const double t0 = .01553241727660705326386197156586357005224;
  const double t1 = .003532809960709087023561817517751309380604;
  const double t2 = .09231719987444221619017126187763868745587;
  const double u0 = .2515001149531479199576969952416196054795;
  const double v0 = -.06287502873828697998942424881040490136987;
  const double P = (double)(1 << d);

  /* set 0: the centre */
  const double w00 = D * (D * (D * a3 + a2) + a1) + a0;
  const double w01 = D * (D * (D * b3 + b2) + b1) - b0;
  w[0][0] = w00;
  w[0][1] = w01;
  w[0][2] = D * c1 - w01;
  w[0][3] = ((D * (D * e2 + e1) - 1.0) + D * (D * (D * a3 + a2) + a1)) + a0;
  w[0][4] = (D * f1 + 1.0) - w00;
  /* set 1: +-lambda1 e_j */
  const double w10 = D * (D * g2 + g1) + g0;
  const double w11 = D * (D * h2 + h1) + h0;
  w[1][0] = w10;
  w[1][1] = w11;
  w[1][2] = k0 - w11;
  w[1][3] = ((D * m1 + m0) + D * (D * g2 + g1)) + g0;
  w[1][4] = n0 - w10;
  /* set 2: +-lambda2 e_j */
  w[2][0] = D * p1 + p0;
  w[2][1] = D * q1 + q0;
  w[2][2] = r0 - (D * q1 + q0);
  w[2][3] = D * s1 + s0;
  w[2][4] = -(D * p1 + p0);
  /* set 3: +-lambda3 e_j */
  w[3][0] = .05769338449097348357291272840392627722165;
  w[3][1] = .03499962660214358382244159694487155861542;
  w[3][2] = -.05769338449097348357291272840392627722165;
  w[3][3] = -1.386862771927828143599782668709014266770;
  w[3][4] = -.05769338449097348357291272840392627722165;
  /* set 4: +-lambda4 e_j */
  w[4][0] = 0.;
  w[4][1] = 0.;
  w[4][2] = -.2386668732575008878964134721962088068396;
  w[4][3] = 0.;
  w[4][4] = 0.;
  /* set 5: (+-lambda1, +-lambda1) */
  w[5][0] = t0 - D * g2;
  w[5][1] = t1 - D * h2;
  w[5][2] = -(t1 - D * h2);
  w[5][3] = (t2 + t0) - D * g2;
  w[5][4] = -(t0 - D * g2);
  /* set 6: (+-lambda1, +-lambda2) */
  w[6][0] = .02254314464717892037990281369120402214829;
  w[6][1] = .01367577326327282236101845043145111753718;
  w[6][2] = -.01367577326327282236101845043145111753718;
  w[6][3] = -.3254475969596012529657378160439011607639;
  w[6][4] = -.02254314464717892037990281369120402214829;
  /* set 7: (+-lambda1, +-lambda1, +-lambda1) */
  w[7][0] = .001770878225839133841300705931694423482268;
  w[7][1] = .001074301277504934385647115949826755327753;
  w[7][2] = -.001074301277504934385647115949826755327753;
  w[7][3] = .001770878225839133841300705931694423482268;
  w[7][4] = -.001770878225839133841300705931694423482268;
  /* set 8: (+-lambda5, ..., +-lambda5) */
  w[8][0] = u0 / P;
  w[8][1] = v0 / P;
  w[8][2] = -(v0 / P);
  w[8][3] = u0 / P;
  w[8][4] = -(u0 / P);
}

static const double kLambda[5] = {
  .4779536579022695061928604197171830064732,
  .2030285873691198677998034402373279133258,
  .4476273546261781288207704806530998539285, .125,
  .3430378987808781457001426145164678603407};

int orc_rule_tables(int d, double* w, double* lambda, double* scale,
                    double* norm, int* count)
{
  if (orc_feval(d) < 0) return -1;
  double W[ORC_NSETS][ORC_NRULES];
  int c[ORC_NSETS];
  rule9_weights(d, W);
  set_counts(d, c);
  for (int s = 0; s < ORC_NSETS; ++s)
    for (int r = 0; r < ORC_NRULES; ++r) {
      if (w) w[s * ORC_NRULES + r] = W[s][r];
      if (scale) scale[s * ORC_NRULES + r] = 0.;
      if (norm) norm[s * ORC_NRULES + r] = 0.;
    }
  /* null-rule scale and norm, Rule.cuh:256-269 (only r = 1..3 defined) */
  for (int s = 0; s < ORC_NSETS; ++s) {
    for (int r = 1; r < ORC_NRULES - 1; ++r) {
      const double sc = (W[s][r] == 0) ? 100 : -W[s][r + 1] / W[s][r];
      double sum = 0;
      for (int x = 0; x < ORC_NSETS; ++x)
        sum += c[x] * fabs(W[x][r + 1] + sc * W[x][r]);
      if (scale) scale[s * ORC_NRULES + r] = sc;
      if (norm) norm[s * ORC_NRULES + r] = 1 / sum;
    }
  }
  if (lambda)
    for (int i = 0; i < 5; ++i) lambda[i] = kLambda[i];
  if (count)
    for (int i = 0; i < ORC_NSETS; ++i) count[i] = c[i];
  return 0;
}

/* Fully symmetric point sets of the degree-9 rule, Rule.cuh:201-249 (the
 * orbit seeds) and :438-505 (enumeration of their signed permutations).
 * The first 1+4d points are ordered centre, (+l1 e0, -l1 e0, +l1 e1, ...),
 * (+l2 e0, -l2 e0, ...), which is the order the fourth difference reads
 * (Sample.cuh:194-218).  The order inside the other sets only changes the
 * summation order. */
int orc_generators(int d, double* gen, int* set)
{
  const int F = orc_feval(d);
  if (F < 0) return -1;
  memset(gen, 0, sizeof(double) * (size_t)F * d);
  int p = 0;
  set[p++] = 0; /* centre */
  for (int s = 1; s <= 4; ++s) {
    for (int j = 0; j < d; ++j) {
      gen[p * d + j] = kLambda[s - 1];
      set[p++] = s;
      gen[p * d + j] = -kLambda[s - 1];
      set[p++] = s;
    }
  }
  /* set 5: (l1, l1) on unordered pairs */
  for (int j = 0; j < d; ++j)
    for (int k = j + 1; k < d; ++k)
      for (int sg = 0; sg < 4; ++sg) {
        gen[p * d + j] = (sg & 1) ? -kLambda[0] : kLambda[0];
        gen[p * d + k] = (sg & 2) ? -kLambda[0] : kLambda[0];
        set[p++] = 5;
      }
  /* set 6: (l1 at j, l2 at k) on ordered pairs */
  for (int j = 0; j < d; ++j)
    for (int k = 0; k < d; ++k) {
      if (k == j) continue;
      for (int sg = 0; sg < 4; ++sg) {
        gen[p * d + j] = (sg & 1) ? -kLambda[0] : kLambda[0];
        gen[p * d + k] = (sg & 2) ? -kLambda[1] : kLambda[1];
        set[p++] = 6;
      }
    }
  /* set 7: (l1, l1, l1) on unordered triples */
  for (int j = 0; j < d; ++j)
    for (int k = j + 1; k < d; ++k)
      for (int l = k + 1; l < d; ++l)
        for (int sg = 0; sg < 8; ++sg) {
          gen[p * d + j] = (sg & 1) ? -kLambda[0] : kLambda[0];
          gen[p * d + k] = (sg & 2) ? -kLambda[0] : kLambda[0];
          gen[p * d + l] = (sg & 4) ? -kLambda[0] : kLambda[0];
          set[p++] = 7;
        }
  /* set 8: (+-l5, ..., +-l5) */
  for (int c = 0; c < (1 << d); ++c) {
    for (int j = 0; j < d; ++j)
      gen[p * d + j] = ((c >> j) & 1) ? -kLambda[4] : kLambda[4];
    set[p++] = 8;
  }
  return p == F ? 0 : -2;
}

/* ======================================================================== */
/* Integrands and closed forms                                               */
/* ======================================================================== */

double orc_integrand(int fam, int d, const double* prm, const double* x)
{
  const double* a = prm;
  const double* u = prm ? prm + d : NULL;
  double s, v;
  switch (fam) {
  case ORC_FAM_OSCILLATORY:
    s = 2.0 * M_PI * u[0];
    for (int i = 0; i < d; ++i) s += a[i] * x[i];
    return cos(s);
  case ORC_FAM_PRODUCT_PEAK:
    v = 1.0;
    for (int i = 0; i < d; ++i) {
      const double t = x[i] - u[i];
      v *= 1.0 / (1.0 / (a[i] * a[i]) + t * t);
    }
    return v;
  case ORC_FAM_CORNER_PEAK:
    s = 1.0;
    for (int i = 0; i < d; ++i) s += a[i] * x[i];
    return pow(s, -(double)(d + 1));
  case ORC_FAM_GAUSSIAN:
    s = 0.0;
    for (int i = 0; i < d; ++i) {
      const double t = x[i] - u[i];
      s += a[i] * a[i] * (t * t);
    }
    return exp(-s);
  case ORC_FAM_C0:
    s = 0.0;
    for (int i = 0; i < d; ++i) s += a[i] * fabs(x[i] - u[i]);
    return exp(-s);
  case ORC_FAM_DISCONTINUOUS:
    for (int i = 0; i < d; ++i)
      if (x[i] > u[i]) return 0.0;
    s = 0.0;
    for (int i = 0; i < d; ++i) s += a[i] * x[i];
    return exp(s);
  case ORC_FAM_CONST:
    return a[0];
  case ORC_FAM_PRODUCT:
    v = 1.0;
    for (int i = 0; i < d; ++i) v *= x[i];
    return v;
  case ORC_FAM_FUN6: {
    /* accuracy_improves_with_epsrel.cu:13-20, arguments (u,v,w,x,y,z) */
    const double nrm = 12.0 / (7.0 - 6 * log(2.0) * log(2.0) + log(64.0));
    return nrm * (x[0] * x[1] + (pow(x[2], x[4]) * x[3] * x[4]) / (1 + x[0]) +
                  x[5] * x[5]);
  }
  case ORC_FAM_BOX:
    s = 0.0;
    for (int i = 0; i < d; ++i) s += x[i] * x[i];
    return pow(s, a[0] / 2);
  case ORC_FAM_NAN:
    return NAN;
  case ORC_FAM_MONOMIAL:
    v = 1.0;
    for (int i = 0; i < d; ++i) v *= pow(x[i], a[i]);
    return v;
  default:
    return NAN;
  }
}

double orc_exact(int fam, int d, const double* prm)
{
  const double* a = prm;
  const double* u = prm ? prm + d : NULL;
  double v = 1.0;
  switch (fam) {
  case ORC_FAM_OSCILLATORY: {
    /* Re[ e^{i 2 pi u_0} prod_j (e^{i a_j} - 1) / (i a_j) ] */
    double re = cos(2.0 * M_PI * u[0]), im = sin(2.0 * M_PI * u[0]);
    for (int j = 0; j < d; ++j) {
      double fr = 1.0, fi = 0.0;
      if (a[j] != 0.0) {
        fr = sin(a[j]) / a[j];
        fi = (1.0 - cos(a[j])) / a[j];
      }
      const double nr = re * fr - im * fi, ni = re * fi + im * fr;
      re = nr;
      im = ni;
    }
    return re;
  }
  case ORC_FAM_PRODUCT_PEAK:
    for (int j = 0; j < d; ++j)
      v *= a[j] * (atan(a[j] * (1.0 - u[j])) + atan(a[j] * u[j]));
    return v;
  case ORC_FAM_CORNER_PEAK: {
    /* inclusion-exclusion over the 2^d corners */
    double fact = 1.0, prod = 1.0, sum = 0.0;
    for (int j = 1; j <= d; ++j) fact *= j;
    for (int j = 0; j < d; ++j) prod *= a[j];
    for (unsigned c = 0; c < (1u << d); ++c) {
      double t = 1.0;
      int bits = 0;
      for (int j = 0; j < d; ++j)
        if ((c >> j) & 1u) {
          t += a[j];
          ++bits;
        }
      sum += ((bits & 1) ? -1.0 : 1.0) / t;
    }
    return sum / (fact * prod);
  }
  case ORC_FAM_GAUSSIAN:
    for (int j = 0; j < d; ++j)
      v *= sqrt(M_PI) / (2.0 * a[j]) *
           (erf(a[j] * (1.0 - u[j])) + erf(a[j] * u[j]));
    return v;
  case ORC_FAM_C0:
    for (int j = 0; j < d; ++j)
      v *= (2.0 - exp(-a[j] * u[j]) - exp(-a[j] * (1.0 - u[j]))) / a[j];
    return v;
  case ORC_FAM_DISCONTINUOUS:
    for (int j = 0; j < d; ++j) {
      const double uu = u[j] < 0.0 ? 0.0 : (u[j] > 1.0 ? 1.0 : u[j]);
      v *= (exp(a[j] * uu) - 1.0) / a[j];
    }
    return v;
  case ORC_FAM_CONST:
    return a[0];
  case ORC_FAM_PRODUCT:
    return ldexp(1.0, -d);
  case ORC_FAM_FUN6:
    return 1.0;
  case ORC_FAM_MONOMIAL:
    for (int j = 0; j < d; ++j) v /= (a[j] + 1.0);
    return v;
  default:
    return NAN;
  }
}

/* ======================================================================== */
/* Per-region evaluation                                                     */
/* ======================================================================== */

typedef struct {
  int d, F;
  double W[ORC_NSETS][ORC_NRULES];
  double scale[ORC_NSETS][ORC_NRULES];
  double norm[ORC_NSETS][ORC_NRULES];
  double* gen;
  int* set;
} rule_t;

static int rule_init(rule_t* R, int d)
{
  R->d = d;
  R->F = orc_feval(d);
  if (R->F < 0) return -1;
  if (orc_rule_tables(d, &R->W[0][0], NULL, &R->scale[0][0], &R->norm[0][0],
                      NULL))
    return -1;
  R->gen = (double*)malloc(sizeof(double) * (size_t)R->F * d);
  R->set = (int*)malloc(sizeof(int) * (size_t)R->F);
  if (!R->gen || !R->set) return -1;
  return orc_generators(d, R->gen, R->set);
}

static void rule_free(rule_t* R)
{
  free(R->gen);
  free(R->set);
}

/* One region: INIT_REGION_POOL (Phases.cuh:301-363) then SampleRegionBlock
 * (Sample.cuh:151-290) with computePermutation (Sample.cuh:103-149). */
static void eval_one(const rule_t* R, int fam, const double* prm,
                     const double* lows, const double* highs,
                     const double* lo, const double* len, double* avg,
                     double* err, int* bisect, double* fd_out, double* rs_out)
{
  const int d = R->d;
  double up[ORC_MAXDIM], range[ORC_MAXDIM], x[ORC_MAXDIM];
  double fv[1 + 4 * ORC_MAXDIM];
  double vol = 1., jac = 1., maxr = 0.;
  int maxdim = 0;
  for (int j = 0; j < d; ++j) {
    up[j] = lo[j] + len[j];
    vol *= up[j] - lo[j];
    range[j] = highs[j] - lows[j];
    jac = jac * range[j];
    const double r = up[j] - lo[j];
    if (r > maxr) {
      maxdim = j;
      maxr = r;
    }
  }
  double s[ORC_NRULES] = {0., 0., 0., 0., 0.};
  double A[ORC_NRULES] = {0., 0., 0., 0., 0.}; /* sum |w_r f|: rounding scale of s_r */
  const int nfd = 1 + 4 * d;
  for (int p = 0; p < R->F; ++p) {
    const double* g = R->gen + (size_t)p * d;
    for (int j = 0; j < d; ++j)
      x[j] = lows[j] + ((.5 + g[j]) * lo[j] + (.5 - g[j]) * up[j]) * range[j];
    const double fun = orc_integrand(fam, d, prm, x) * jac;
    if (p < nfd) fv[p] = fun;
    const double* w = R->W[R->set[p]];
    for (int r = 0; r < ORC_NRULES; ++r) {
      s[r] += fun * w[r];
      A[r] += fabs(fun * w[r]);
    }
  }
  /* fourth difference and split axis, Sample.cuh:194-218 */
  const double l1 = kLambda[0], l2 = kLambda[1];
  const double q = l2 / l1;
  const double ratio = q * q;
  const double base = fv[0] * 2 * (1 - ratio);
  double maxdiff = 0.;
  double fdmag = 0.; /* magnitude of the terms of the largest difference */
  int bis = maxdim;
  for (int j = 0; j < d; ++j) {
    const double fp1 = fv[1 + 2 * j], fm1 = fv[2 + 2 * j];
    const double fp2 = fv[1 + 2 * d + 2 * j], fm2 = fv[2 + 2 * d + 2 * j];
    const double fd = fabs(base + ratio * (fp1 + fm1) - (fp2 + fm2));
    const double mag = fabs(base) + ratio * (fabs(fp1) + fabs(fm1)) + fabs(fp2) + fabs(fm2);
    if (mag > fdmag) fdmag = mag;
    if (fd_out) fd_out[j] = fd;
    if (fd > maxdiff) {
      maxdiff = fd;
      bis = j;
    }
  }
  /* null-rule error, Sample.cuh:261-289 (CUDA max == fmax) */
  double e[ORC_NRULES];
  for (int r = 1; r < ORC_NRULES - 1; ++r) {
    double me = 0.;
    for (int k = 0; k < ORC_NSETS; ++k)
      me = fmax(me, fabs(s[r + 1] + R->scale[k][r] * s[r]) * R->norm[k][r]);
    e[r] = me;
  }
  if (rs_out) {
    for (int r = 0; r < ORC_NRULES; ++r) {
      rs_out[r] = s[r];
      rs_out[ORC_NRULES + r] = A[r];
    }
    rs_out[10] = vol;
    rs_out[11] = e[1];
    rs_out[12] = e[2];
    rs_out[13] = e[3];
    rs_out[14] = fdmag;
    rs_out[15] = 0.0;
  }
  *avg = vol * s[0];
  *err = vol * ((5. * e[1] <= e[2] && 5. * e[2] <= e[3])
                  ? 1. * e[1]
                  : 5. * fmax(fmax(e[1], e[2]), e[3]));
  *bisect = bis;
}

int orc_eval_regions(int d, int fam, const double* prm, const double* lows,
                     const double* highs, size_t n, const double* L,
                     const double* H, double* est, double* err, int* dim,
                     double* fdiff, double* rsum)
{
  rule_t R;
  if (rule_init(&R, d)) {
    rule_free(&R);
    return -1;
  }
  double lw[ORC_MAXDIM], hg[ORC_MAXDIM];
  for (int j = 0; j < d; ++j) {
    lw[j] = lows ? lows[j] : 0.0;
    hg[j] = highs ? highs[j] : 1.0;
  }
  const long long nn = (long long)n;
#pragma omp parallel for schedule(static) if (nn > 64)
  for (long long t = 0; t < nn; ++t) {
    double lo[ORC_MAXDIM], len[ORC_MAXDIM];
    for (int j = 0; j < d; ++j) {
      lo[j] = L[(size_t)j * n + (size_t)t];
      len[j] = H[(size_t)j * n + (size_t)t];
    }
    eval_one(&R, fam, prm, lw, hg, lo, len, &est[t], &err[t], &dim[t],
             fdiff ? fdiff + (size_t)t * d : NULL,
             rsum ? rsum + (size_t)t * ORC_RSUM : NULL);
  }
  rule_free(&R);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Evaluation at the library kernel's own points, in extended precision.     */
/*                                                                          */
/* The restatement above maps every point as the reference does,            */
/* x = lows + ((.5 + g) lo + (.5 - g) up) range (Sample.cuh:124-131).  The   */
/* library kernel forms its coordinates differently (eval_kernel.hpp,        */
/* eval_regions_wave): the region's centre c = fma(.5 (lo + up), range, lows) */
/* and width h = (up - lo) range, minus the integrand's origin and times     */
/* its axis scale when it declares them, and every point v = fma(-g, h, c). */
/* Those are a few roundings away from the reference's x, and an integrand   */
/* with a steep exponent (a Gaussian with a = 25) turns them into relative   */
/* differences of 1e-12 in f.  This mode forms the points exactly as the     */
/* kernel does (in double, bit for bit), and evaluates f at them and every  */
/* rule sum in long double, so what is left of a difference from the kernel */
/* is the kernel's own rounding of f and of its sums.                        */
/* ------------------------------------------------------------------------ */

/* origin and axis scale the library's Genz functors declare
 * (include/pagani/genz.hpp); plain integrands have neither. */
static void kform_shift(int fam, int d, const double* prm, int plain, double* org, double* scl)
{
  const double* a = prm;
  const double* u = prm ? prm + d : NULL;
  for (int j = 0; j < d; ++j) {
    org[j] = 0.0;
    scl[j] = 1.0;
  }
  if (plain) return;
  switch (fam) {
  case ORC_FAM_OSCILLATORY:
  case ORC_FAM_CORNER_PEAK:
    for (int j = 0; j < d; ++j) scl[j] = a[j];
    break;
  case ORC_FAM_PRODUCT_PEAK:
  case ORC_FAM_GAUSSIAN:
  case ORC_FAM_C0:
    for (int j = 0; j < d; ++j) {
      org[j] = u[j];
      scl[j] = a[j];
    }
    break;
  case ORC_FAM_DISCONTINUOUS:
    for (int j = 0; j < d; ++j) org[j] = u[j];
    break;
  default:
    break;
  }
}

/* f at kernel coordinates v (centred and scaled Genz, or plain x), in long
 * double.  The constants the library's functors hold in double (the
 * oscillatory phase 2 pi u_0, the product peak's prod a_i^2, the
 * discontinuous family's sum a_i u_i) are formed in double as they are. */
/* *cond receives f's sensitivity to one rounding of each operation of its
 * inner sum: |f| times the sum of the magnitudes that sum adds up (|sin|
 * times them for the cosine), the error any FP64 evaluation of f carries in
 * units of 2^-53. */
static long double kform_f(int fam, int d, const double* prm, int plain, const double* v,
                           long double* cond)
{
  const double* a = prm;
  const double* u = prm ? prm + d : NULL;
  long double s = 0.0L, p = 1.0L, m = 0.0L, f;
  if (plain) {
    switch (fam) {
    case ORC_FAM_OSCILLATORY:
      s = 2.0L * 3.14159265358979323846264338327950288L * u[0];
      m = fabsl(s);
      for (int i = 0; i < d; ++i) {
        s += (long double)a[i] * v[i];
        m += fabsl((long double)a[i] * v[i]);
      }
      *cond = fabsl(sinl(s)) * m;
      return cosl(s);
    case ORC_FAM_PRODUCT_PEAK:
      for (int i = 0; i < d; ++i) {
        const long double t = (long double)v[i] - u[i];
        p *= 1.0L / (1.0L / ((long double)a[i] * a[i]) + t * t);
        m += 4.0L + 2.0L * fabsl(t) * fmaxl(fabsl(v[i]), fabsl(u[i])) /
                        (1.0L / ((long double)a[i] * a[i]) + t * t);
      }
      *cond = fabsl(p) * m;
      return p;
    case ORC_FAM_CORNER_PEAK:
      s = 1.0L;
      m = 1.0L;
      for (int i = 0; i < d; ++i) {
        s += (long double)a[i] * v[i];
        m += fabsl((long double)a[i] * v[i]);
      }
      f = powl(s, -(long double)(d + 1));
      *cond = fabsl(f) * (d + 1) * (1.0L + m / fabsl(s));
      return f;
    case ORC_FAM_GAUSSIAN:
      for (int i = 0; i < d; ++i) {
        const long double t = (long double)v[i] - u[i];
        s += (long double)a[i] * a[i] * t * t;
        m += (long double)a[i] * a[i] * fabsl(t) * (fabsl(t) + 2 * fmaxl(fabsl(v[i]), fabsl(u[i])));
      }
      f = expl(-s);
      *cond = f * (m + 1.0L);
      return f;
    case ORC_FAM_C0:
      for (int i = 0; i < d; ++i) {
        const long double t = fabsl((long double)v[i] - u[i]);
        s += (long double)a[i] * t;
        m += (long double)a[i] * (t + fmaxl(fabsl(v[i]), fabsl(u[i])));
      }
      f = expl(-s);
      *cond = f * (m + 1.0L);
      return f;
    case ORC_FAM_DISCONTINUOUS:
      *cond = 0.0L;
      for (int i = 0; i < d; ++i)
        if (v[i] > u[i]) return 0.0L;
      for (int i = 0; i < d; ++i) {
        s += (long double)a[i] * v[i];
        m += fabsl((long double)a[i] * v[i]);
      }
      f = expl(s);
      *cond = f * (m + 1.0L);
      return f;
    default:
      f = orc_integrand(fam, d, prm, v);
      *cond = fabsl(f) * d;
      return f;
    }
  }
  switch (fam) {
  case ORC_FAM_OSCILLATORY:
    s = (long double)(2.0 * M_PI * u[0]);
    m = fabsl(s);
    for (int i = 0; i < d; ++i) {
      s += v[i];
      m += fabsl((long double)v[i]);
    }
    *cond = fabsl(sinl(s)) * m;
    return cosl(s);
  case ORC_FAM_PRODUCT_PEAK: {
    double k2 = 1.0;
    for (int i = 0; i < d; ++i) k2 *= a[i] * a[i];
    for (int i = 0; i < d; ++i) p *= 1.0L + (long double)v[i] * v[i];
    f = (long double)k2 / p;
    *cond = f * (3.0L * d + 1.0L);
    return f;
  }
  case ORC_FAM_CORNER_PEAK:
    s = 1.0L;
    m = 1.0L;
    for (int i = 0; i < d; ++i) {
      s += v[i];
      m += fabsl((long double)v[i]);
    }
    f = powl(s, -(long double)(d + 1));
    *cond = fabsl(f) * (d + 1) * (1.0L + m / fabsl(s));
    return f;
  case ORC_FAM_GAUSSIAN:
    for (int i = 0; i < d; ++i) s += (long double)v[i] * v[i];
    f = expl(-s);
    *cond = f * (s + 1.0L);
    return f;
  case ORC_FAM_C0:
    for (int i = 0; i < d; ++i) s += fabsl((long double)v[i]);
    f = expl(-s);
    *cond = f * (s + 1.0L);
    return f;
  case ORC_FAM_DISCONTINUOUS: {
    double c = 0.0;
    for (int i = 0; i < d; ++i) c += a[i] * u[i];
    *cond = 0.0L;
    for (int i = 0; i < d; ++i)
      if (v[i] > 0.0) return 0.0L;
    s = c;
    m = fabsl((long double)c);
    for (int i = 0; i < d; ++i) {
      s += (long double)a[i] * v[i];
      m += fabsl((long double)a[i] * v[i]);
    }
    f = expl(s);
    *cond = f * (m + 1.0L);
    return f;
  }
  default:
    f = orc_integrand(fam, d, prm, v);
    *cond = fabsl(f) * d;
    return f;
  }
}

static void eval_one_kform(const rule_t* R, int fam, const double* prm, int plain,
                           const double* lows, const double* highs, const double* lo,
                           const double* len, double* avg, double* err, int* bisect,
                           double* fd_out, double* rs_out)
{
  const int d = R->d;
  double c[ORC_MAXDIM], h[ORC_MAXDIM], org[ORC_MAXDIM], scl[ORC_MAXDIM], v[ORC_MAXDIM];
  long double fv[1 + 4 * ORC_MAXDIM];
  double vol = 1., jac = 1., maxr = 0.;
  int maxdim = 0;
  kform_shift(fam, d, prm, plain, org, scl);
  for (int j = 0; j < d; ++j) {
    const double up = lo[j] + len[j];
    const double r = up - lo[j];
    const double range = highs[j] - lows[j];
    vol *= r;
    jac = jac * range;
    if (r > maxr) {
      maxdim = j;
      maxr = r;
    }
    c[j] = fma(0.5 * (lo[j] + up), range, lows[j]);
    h[j] = r * range;
    if (!plain) {
      c[j] -= org[j];
      c[j] *= scl[j];
      h[j] *= scl[j];
    }
  }
  long double s[ORC_NRULES] = {0, 0, 0, 0, 0};
  long double A[ORC_NRULES] = {0, 0, 0, 0, 0};
  long double C = 0.0L;
  const int nfd = 1 + 4 * d;
  for (int p = 0; p < R->F; ++p) {
    const double* g = R->gen + (size_t)p * d;
    for (int j = 0; j < d; ++j) v[j] = fma(-g[j], h[j], c[j]);
    long double cp = 0.0L;
    const long double fun = kform_f(fam, d, prm, plain, v, &cp) * jac;
    C += fabsl(cp * jac * R->W[R->set[p]][0]);
    if (p < nfd) fv[p] = fun;
    const double* w = R->W[R->set[p]];
    for (int r = 0; r < ORC_NRULES; ++r) {
      s[r] += fun * w[r];
      A[r] += fabsl(fun * w[r]);
    }
  }
  const long double l1 = kLambda[0], l2 = kLambda[1];
  const long double q = l2 / l1;
  const long double ratio = q * q;
  const long double base = fv[0] * 2 * (1 - ratio);
  long double maxdiff = 0.;
  long double fdmag = 0.;
  int bis = maxdim;
  for (int j = 0; j < d; ++j) {
    const long double fp1 = fv[1 + 2 * j], fm1 = fv[2 + 2 * j];
    const long double fp2 = fv[1 + 2 * d + 2 * j], fm2 = fv[2 + 2 * d + 2 * j];
    const long double fd = fabsl(base + ratio * (fp1 + fm1) - (fp2 + fm2));
    const long double mag = fabsl(base) + ratio * (fabsl(fp1) + fabsl(fm1)) + fabsl(fp2) +
                            fabsl(fm2);
    if (mag > fdmag) fdmag = mag;
    if (fd_out) fd_out[j] = (double)fd;
    if (fd > maxdiff) {
      maxdiff = fd;
      bis = j;
    }
  }
  long double e[ORC_NRULES];
  for (int r = 1; r < ORC_NRULES - 1; ++r) {
    long double me = 0.;
    for (int k = 0; k < ORC_NSETS; ++k)
      me = fmaxl(me, fabsl(s[r + 1] + R->scale[k][r] * s[r]) * R->norm[k][r]);
    e[r] = me;
  }
  if (rs_out) {
    for (int r = 0; r < ORC_NRULES; ++r) {
      rs_out[r] = (double)s[r];
      rs_out[ORC_NRULES + r] = (double)A[r];
    }
    rs_out[10] = vol;
    rs_out[11] = (double)e[1];
    rs_out[12] = (double)e[2];
    rs_out[13] = (double)e[3];
    rs_out[14] = (double)fdmag;
    rs_out[15] = (double)C;
  }
  *avg = (double)(vol * s[0]);
  *err = (double)(vol * ((5. * e[1] <= e[2] && 5. * e[2] <= e[3])
                             ? 1. * e[1]
                             : 5. * fmaxl(fmaxl(e[1], e[2]), e[3])));
  *bisect = bis;
}

int orc_eval_regions_kform(int d, int fam, const double* prm, int plain, const double* lows,
                           const double* highs, size_t n, const double* L, const double* H,
                           double* est, double* err, int* dim, double* fdiff, double* rsum)
{
  rule_t R;
  if (rule_init(&R, d)) {
    rule_free(&R);
    return -1;
  }
  double lw[ORC_MAXDIM], hg[ORC_MAXDIM];
  for (int j = 0; j < d; ++j) {
    lw[j] = lows ? lows[j] : 0.0;
    hg[j] = highs ? highs[j] : 1.0;
  }
  const long long nn = (long long)n;
#pragma omp parallel for schedule(static) if (nn > 64)
  for (long long t = 0; t < nn; ++t) {
    double lo[ORC_MAXDIM], len[ORC_MAXDIM];
    for (int j = 0; j < d; ++j) {
      lo[j] = L[(size_t)j * n + (size_t)t];
      len[j] = H[(size_t)j * n + (size_t)t];
    }
    eval_one_kform(&R, fam, prm, plain, lw, hg, lo, len, &est[t], &err[t], &dim[t],
                   fdiff ? fdiff + (size_t)t * d : NULL, rsum ? rsum + (size_t)t * ORC_RSUM : NULL);
  }
  rule_free(&R);
  return 0;
}

/* ======================================================================== */
/* Region-list phases                                                        */
/* ======================================================================== */

/* create_uniform_split, Sub_regions.cuh:11-36 and :165-185 */
void orc_uniform_split(int d, size_t k, double* L, double* H)
{
  size_t n = 1;
  for (int j = 0; j < d; ++j) n *= k;
  const double len = 1. / (double)k;
  for (size_t t = 0; t < n; ++t) {
    size_t q = t;
    for (int j = 0; j < d; ++j) {
      const size_t id = q % k;
      q /= k;
      L[(size_t)j * n + t] = (double)id * len;
      H[(size_t)j * n + t] = len;
    }
  }
}

/* RefineError, Phases.cuh:219-268, with MaxErr from cudaUtil.h:26-30.
 * Children t and t+n/2 share parent t mod n/2. */
void orc_refine(size_t n, const double* est, double* err,
                const double* pest, double epsrel, int relerr,
                uint8_t* active)
{
  const size_t h = n / 2;
  for (size_t i = 0; i < h; ++i) {
    const size_t l = i, r = i + h;
    const double el = err[l], er = err[r];
    const double diff = fabs(.25 * (est[r] + est[l] - pest[i]));
    const double E = el + er;
    double nl = el, nr = er;
    if (E > 0.0) {
      const double c = 1 + 2 * diff / E;
      nl *= c;
      nr *= c;
    }
    nl += diff;
    nr += diff;
    err[l] = nl;
    err[r] = nr;
    const int pl = relerr && nl < fmax(epsrel * fabs(est[l]), 1e-200);
    const int pr = relerr && nr < fmax(epsrel * fabs(est[r]), 1e-200);
    active[l] = (uint8_t)!pl;
    active[r] = (uint8_t)!pr;
  }
}

/* get_num_active_regions + alignRegions, Sub_region_filter.cuh:8-47, 62-139 */
size_t orc_filter(int d, size_t n, const double* L, const double* H,
                  const uint8_t* active, const double* est, const double* err,
                  const int* dim, double* L2, double* H2, double* pest,
                  double* perr, int* dim2)
{
  size_t m = 0;
  for (size_t t = 0; t < n; ++t) m += active[t] == 1;
  if (m == 0) return 0;
  size_t idx = 0;
  for (size_t t = 0; t < n; ++t) {
    if (active[t] != 1) continue;
    for (int j = 0; j < d; ++j) {
      L2[(size_t)j * m + idx] = L[(size_t)j * n + t];
      H2[(size_t)j * m + idx] = H[(size_t)j * n + t];
    }
    pest[idx] = est[t];
    perr[idx] = err[t];
    dim2[idx] = dim[t];
    ++idx;
  }
  return m;
}

/* divideIntervalsGPU, Sub_region_splitter.cuh:8-46 */
void orc_split(int d, size_t m, const double* L, const double* H,
               const int* dim, double* L2, double* H2)
{
  const size_t n2 = 2 * m;
  for (size_t i = 0; i < m; ++i) {
    for (int c = 0; c < 2; ++c)
      for (int j = 0; j < d; ++j) {
        L2[(size_t)j * n2 + c * m + i] = L[(size_t)j * m + i];
        H2[(size_t)j * n2 + c * m + i] = H[(size_t)j * m + i];
      }
    const int b = dim[i];
    const double half = H[(size_t)b * m + i] / 2;
    for (int c = 0; c < 2; ++c) {
      L2[(size_t)b * n2 + c * m + i] = L[(size_t)b * m + i] + c * half;
      H2[(size_t)b * n2 + c * m + i] = half;
    }
  }
}

/* ======================================================================== */
/* Heuristic classifier, heuristic_classifier.cuh:89-452                     */
/* ======================================================================== */

void orc_clf_init(orc_classifier* c, int d, double epsrel, double epsabs,
                  size_t total_mem)
{
  memset(c, 0, sizeof(*c));
  c->epsrel = epsrel;
  c->epsabs = epsabs;
  c->required_digits = (int)ceil(log10(1 / epsrel)); /* :164-167 */
  c->max_budget = .25;
  c->max_active = .5;
  c->total_mem = total_mem;
  c->ndim = d;
  /* hist[] is left uninitialised in the reference (:153); zero here */
}

void orc_clf_store(orc_classifier* c, double estimate)
{
  c->hist[0] = c->hist[1];
  c->hist[1] = c->hist[2];
  c->hist[2] = estimate;
  c->iters_collected++;
}

/* sigDigitsSame, :170-203: compare the leading decimal characters of the
 * last three estimates printed in std::fixed with precision 15. */
int orc_clf_sig_digits_same(const orc_classifier* c)
{
  double third = fabs(c->hist[0]);
  double second = fabs(c->hist[1]);
  double first = fabs(c->hist[2]);
  while (first != 0. && first < 1.) first *= 10;
  while (second != 0. && second < 1.) second *= 10;
  while (third != 0. && third < 1.) third *= 10;
  char s2l[512], last[512], cur[512];
  memset(s2l, 0, sizeof s2l);
  memset(last, 0, sizeof last);
  memset(cur, 0, sizeof cur);
  snprintf(s2l, sizeof s2l - 1, "%.15f", third);
  snprintf(last, sizeof last - 1, "%.15f", second);
  snprintf(cur, sizeof cur - 1, "%.15f", first);
  int verdict = 1, sig = 0;
  const int rd = c->required_digits;
  for (int i = 0; i < rd + 1 && sig < rd && verdict == 1 && i < 500; ++i) {
    verdict = (cur[i] == last[i] && last[i] == s2l[i]) ? 1 : 0;
    sig += (verdict == 1 && cur[i] != '.') ? 1 : 0;
  }
  return verdict;
}

/* estimate_converged, :205-216 (min_iters_for_convergence = 1) */
int orc_clf_estimate_converged(const orc_classifier* c)
{
  if (c->iters_collected - 1 < 1 || !orc_clf_sig_digits_same(c)) return 0;
  return 1;
}

/* num_doubles_needed / num_ints_needed / device_mem_required_for_full_split,
 * :96-133 */
size_t orc_mem_required(size_t n, int d)
{
  const size_t nd = n * (size_t)d;
  const size_t dbl = n + n + 2 * n + 2 * n + 2 * nd + 2 * nd + 2 * nd +
                     2 * nd + n + n + nd + nd;
  const size_t ints = n + 2 * n + n;
  return 8 * dbl + 4 * ints;
}

/* free_device_mem, :135-145 (total_device_mem() is 16e9 in the reference;
 * here the total is a parameter) */
size_t orc_free_mem(size_t n, int d, size_t total)
{
  const size_t occ = orc_mem_required(n, d);
  return total > occ ? total - occ : 1;
}

static double mem_ratio(size_t n, int d, size_t total)
{
  return (double)orc_mem_required(n, d) / (double)orc_free_mem(n, d, total);
}

/* classification_criteria_met, :348-360 */
int orc_clf_criteria_met(const orc_classifier* c, size_t n)
{
  const double ratio = mem_ratio(n, c->ndim, c->total_mem);
  if (ratio > 1.) return 1;
  if (ratio > .1 && orc_clf_estimate_converged(c)) return 1;
  return 0;
}

/* apply_threshold, :282-303 */
static void apply_threshold(const orc_classifier* c, orc_class_res* res,
                            const double* err, size_t n, uint8_t* flags)
{
  size_t cnt = 0;
  for (size_t i = 0; i < n; ++i) {
    flags[i] = err[i] > res->threshold;
    cnt += flags[i];
  }
  res->threshold_passes++;
  res->num_active = cnt;
  res->percent_mem_active = (double)(int)cnt / (double)(int)n;
  res->pass_mem = res->percent_mem_active <= c->max_active;
}

/* get_larger_threshold_results, :327-346 */
static void larger_threshold(const orc_classifier* c, orc_class_res* res,
                             const double* err, size_t n, uint8_t* flags)
{
  res->pass_mem = 0;
  size_t counter = 0;
  while (!res->pass_mem && counter < 20) {
    apply_threshold(c, res, err, n, flags);
    if (!res->pass_mem) {
      res->range_low = res->threshold;
      const double diff = fabs(res->range_high - res->threshold);
      res->threshold += diff * .5; /* increase_threshold, :37-42 */
    }
    counter++;
  }
}

/* classify, :362-451 */
void orc_clf_classify(orc_classifier* c, const double* err, size_t n,
                      double iter_errorest, double iter_finished_errorest,
                      double total_finished_errorest, uint8_t* flags,
                      orc_class_res* res)
{
  memset(res, 0, sizeof(*res));
  double mn = err[0], mx = err[0];
  for (size_t i = 1; i < n; ++i) {
    if (err[i] < mn) mn = err[i];
    if (mx < err[i]) mx = err[i];
  }
  res->range_low = mn;
  res->range_high = mx;
  res->threshold = iter_errorest / (double)n;
  res->max_budget_perc_to_cover = .25;
  res->max_active_perc = .5;
  const double target_error = fabs(c->hist[2]) * c->epsrel;
  size_t n_inc = 0, n_dec = 0;
  do {
    if (!res->pass_mem && n_inc <= 20) {
      larger_threshold(c, res, err, n, flags);
      n_inc++;
    }
    if (res->pass_mem) {
      /* evaluate_error_budget, :305-325 */
      double dot = 0.;
      for (size_t i = 0; i < n; ++i) dot += err[i] * (double)flags[i];
      const double extra = iter_errorest - dot - iter_finished_errorest;
      const double budget = target_error - total_finished_errorest;
      res->pass_errorest_budget = extra <= c->max_budget * budget;
      res->finished_errorest = extra;
      if (!res->pass_errorest_budget && n_dec <= 20) {
        res->range_high = res->threshold;
        const double diff = fabs(res->range_low - res->threshold);
        res->threshold -= diff * .5; /* decrease_threshold, :30-35 */
        res->pass_mem = 0;
        n_dec++;
      }
    }
    const int exhausted = n_dec >= 20 || n_inc >= 20;
    if (exhausted && c->max_budget < .7) {
      c->max_budget += 0.1;
      n_dec = 0;
      n_inc = 0;
      res->range_low = mn;
      res->range_high = mx;
      res->threshold = iter_errorest / (double)n;
    } else if (exhausted && c->max_budget >= .7 && c->max_active <= .7) {
      c->max_active += .1;
    } else if (exhausted) {
      break;
    }
  } while (!res->pass_mem || !res->pass_errorest_budget);
  res->max_budget_perc_to_cover = c->max_budget;
  res->max_active_perc = c->max_active;
  c->max_active = .5;
  c->max_budget = .25;
}

/* ======================================================================== */
/* The iteration loop, Workspace.cuh:362-494 (and :151-358 for the          */
/* Sub_regions& overload's bookkeeping)                                      */
/* ======================================================================== */

static double seq_sum(const double* a, size_t n)
{
  double s = 0.;
  for (size_t i = 0; i < n; ++i) s += a[i];
  return s;
}

static double seq_dot(const uint8_t* f, const double* a, size_t n)
{
  double s = 0.;
  for (size_t i = 0; i < n; ++i) s += (double)f[i] * a[i];
  return s;
}

static void trace_push(orc_trace* tr, size_t n, size_t m, double e, double r)
{
  if (!tr || tr->len >= tr->cap) return;
  tr->n[tr->len] = n;
  tr->m[tr->len] = m;
  tr->estimate[tr->len] = e;
  tr->errorest[tr->len] = r;
  tr->len++;
}

int orc_integrate(int d, int fam, const double* prm, double epsrel,
                  double epsabs, const double* lows_in,
                  const double* highs_in, int relerr, size_t total_mem,
                  size_t init_k, int user_regions, orc_result* out,
                  orc_trace* tr)
{
  return orc_integrate_ex(d, fam, prm, epsrel, epsabs, lows_in, highs_in, relerr, total_mem,
                          init_k, user_regions, out, tr, NULL);
}

void orc_free(void* p) { free(p); }

/* The final list goes to the options (ownership passes to the caller). */
static void keep_list(orc_opts* o, int d, size_t n, double* L, double* H, double* est,
                      double* err, const int* dim)
{
  o->list_n = n;
  o->list_L = (double*)malloc(sizeof(double) * (n ? n * d : 1));
  o->list_H = (double*)malloc(sizeof(double) * (n ? n * d : 1));
  memcpy(o->list_L, L, sizeof(double) * n * d);
  memcpy(o->list_H, H, sizeof(double) * n * d);
  o->list_est = o->list_err = NULL;
  o->list_dim = NULL;
  if (est) {
    o->list_est = (double*)malloc(sizeof(double) * (n ? n : 1));
    o->list_err = (double*)malloc(sizeof(double) * (n ? n : 1));
    o->list_dim = (int*)malloc(sizeof(int) * (n ? n : 1));
    memcpy(o->list_est, est, sizeof(double) * n);
    memcpy(o->list_err, err, sizeof(double) * n);
    memcpy(o->list_dim, dim, sizeof(int) * n);
  }
}

int orc_integrate_ex(int d, int fam, const double* prm, double epsrel,
                     double epsabs, const double* lows_in,
                     const double* highs_in, int relerr, size_t total_mem,
                     size_t init_k, int user_regions, orc_result* out,
                     orc_trace* tr, orc_opts* opts)
{
  const int keep = opts && opts->keep_list;
  const int predict_split = opts && opts->predict_split;
  const size_t max_iters = opts ? opts->max_iters : 0;
  if (opts) {
    opts->list_n = 0;
    opts->list_L = opts->list_H = opts->list_est = opts->list_err = NULL;
    opts->list_dim = NULL;
    opts->budget_fixes = 0;
  }
  if (orc_feval(d) < 0 || !out) return -1;
  double lows[ORC_MAXDIM], highs[ORC_MAXDIM];
  for (int j = 0; j < d; ++j) {
    lows[j] = lows_in ? lows_in[j] : 0.0;
    highs[j] = highs_in ? highs_in[j] : 1.0;
  }
  if (total_mem == 0) total_mem = (size_t)16e9;
  size_t k = init_k;
  if (k == 0) k = d < 5 ? 4 : (d <= 10 ? 2 : 1); /* :379-385 */
  size_t n = 1;
  for (int j = 0; j < d; ++j) n *= k;

  double* L = (double*)malloc(sizeof(double) * n * d);
  double* H = (double*)malloc(sizeof(double) * n * d);
  double *pest = NULL, *perr = NULL;
  size_t np = 0;
  orc_uniform_split(d, k, L, H);

  orc_result cum;
  memset(&cum, 0, sizeof cum);
  cum.status = 1;
  cum.lastPhase = -1;
  orc_classifier clf;
  orc_clf_init(&clf, d, epsrel, epsabs, total_mem);
  int rc = 0;

  for (size_t it = 0; it < 700 && n > 0; it++) {
    if (max_iters && it == max_iters) {
      if (keep) keep_list(opts, d, n, L, H, NULL, NULL, NULL);
      rc = 2;
      goto done;
    }
    double* est = (double*)malloc(sizeof(double) * n);
    double* err = (double*)malloc(sizeof(double) * n);
    int* dim = (int*)malloc(sizeof(int) * n);
    uint8_t* act = (uint8_t*)malloc(n);
    orc_eval_regions(d, fam, prm, lows, highs, n, L, H, est, err, dim, NULL,
                     NULL);
    memset(act, 1, n);
    const double iter_est = seq_sum(est, n);
    /* predict_split overrides the relerr flag per iteration,
     * Workspace.cuh:206-210 / 415-420 */
    const int relerr_it =
        predict_split ? ((n <= 15000000 && it < 15 && cum.nregions == 0) ? 0 : 1) : relerr;
    if (np != 0) orc_refine(n, est, err, pest, epsrel, relerr_it, act);
    const double iter_err = seq_sum(err, n);
    if (user_regions) cum.iters++;

    const double tot_est = fabs(cum.estimate + iter_est);
    const double tot_err = cum.errorest + iter_err;
    /* accuracy_reached, PaganiUtils.cuh:387-394 */
    if (tot_err / tot_est <= epsrel || tot_err <= epsabs) {
      trace_push(tr, n, 0, cum.estimate + iter_est, tot_err);
      cum.estimate += iter_est;
      cum.errorest += iter_err;
      cum.status = 0;
      cum.nregions += n;
      if (keep) keep_list(opts, d, n, L, H, est, err, dim);
      free(est), free(err), free(dim), free(act);
      goto done;
    }
    orc_clf_store(&clf, cum.estimate + iter_est);
    /* compute_finished_estimates, PaganiUtils.cuh:368-385 */
    double fin_est = iter_est - seq_dot(act, est, n);
    double fin_err = iter_err - seq_dot(act, err, n);
    /* fix_error_budget_overflow, Workspace.cuh:121-146 */
    if (cum.errorest + fin_err > fabs(cum.estimate + iter_est) * epsrel) {
      if (opts) opts->budget_fixes++;
      memset(act, 1, n);
      fin_err = 0.;
      fin_est = 0.;
    }
    /* heuristic_classify, Workspace.cuh:74-118 */
    {
      const double ratio = mem_ratio(n, d, total_mem);
      const int necessary = ratio > 1.;
      int terminate;
      if (!orc_clf_criteria_met(&clf, n)) {
        terminate = necessary;
      } else {
        uint8_t* fl = (uint8_t*)malloc(n);
        orc_class_res res;
        orc_clf_classify(&clf, err, n, iter_err, fin_err, cum.errorest, fl,
                         &res);
        const int ok = res.pass_mem && res.pass_errorest_budget;
        if (ok) {
          memcpy(act, fl, n);
          fin_est = iter_est - seq_dot(act, est, n);
          fin_err = res.finished_errorest;
        }
        terminate = (!ok && necessary) || res.num_active == 0;
        free(fl);
      }
      if (terminate) {
        trace_push(tr, n, 0, cum.estimate + iter_est, tot_err);
        cum.estimate += iter_est;
        cum.errorest += iter_err;
        cum.nregions += n;
        if (keep) keep_list(opts, d, n, L, H, est, err, dim);
        free(est), free(err), free(dim), free(act);
        goto done;
      }
    }
    /* Sub_regions_filter::filter then Sub_region_splitter::split */
    const double rec_est = cum.estimate + iter_est;
    size_t m = 0;
    for (size_t t = 0; t < n; ++t) m += act[t] == 1;
    trace_push(tr, n, m, rec_est, tot_err);
    cum.estimate += fin_est;
    cum.errorest += fin_err;
    cum.nregions += n - m;
    if (m > 0) {
      double* L2 = (double*)malloc(sizeof(double) * m * d);
      double* H2 = (double*)malloc(sizeof(double) * m * d);
      int* dim2 = (int*)malloc(sizeof(int) * m);
      free(pest);
      free(perr);
      pest = (double*)malloc(sizeof(double) * m);
      perr = (double*)malloc(sizeof(double) * m);
      orc_filter(d, n, L, H, act, est, err, dim, L2, H2, pest, perr, dim2);
      np = m;
      free(L);
      free(H);
      L = (double*)malloc(sizeof(double) * 2 * m * d);
      H = (double*)malloc(sizeof(double) * 2 * m * d);
      orc_split(d, m, L2, H2, dim2, L, H);
      free(L2), free(H2), free(dim2);
      n = 2 * m;
    } else {
      n = 0;
    }
    free(est), free(err), free(dim), free(act);
    if (!user_regions) cum.iters++;
  }
  cum.nregions += n;
  if (keep) keep_list(opts, d, n, L, H, NULL, NULL, NULL);
done:
  free(L);
  free(H);
  free(pest);
  free(perr);
  *out = cum;
  return rc;
}

/* ======================================================================== */
/* Sequential Cuhre-equivalent (CPU baseline)                                */
/* ======================================================================== */

typedef struct {
  double avg, err;
  int bisect;
  double lo[ORC_MAXDIM], len[ORC_MAXDIM];
} creg_t;

static void heap_push(creg_t* h, size_t* n, const creg_t* r)
{
  size_t i = (*n)++;
  h[i] = *r;
  while (i > 0) {
    const size_t p = (i - 1) / 2;
    if (h[p].err >= h[i].err) break;
    creg_t t = h[p];
    h[p] = h[i];
    h[i] = t;
    i = p;
  }
}

static creg_t heap_pop(creg_t* h, size_t* n)
{
  creg_t top = h[0];
  h[0] = h[--(*n)];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, b = i;
    if (l < *n && h[l].err > h[b].err) b = l;
    if (r < *n && h[r].err > h[b].err) b = r;
    if (b == i) break;
    creg_t t = h[b];
    h[b] = h[i];
    h[i] = t;
    i = b;
  }
  return top;
}

int orc_cuhre_seq(int d, int fam, const double* prm, double epsrel,
                  double epsabs, size_t maxeval, orc_result* out)
{
  rule_t R;
  if (rule_init(&R, d)) {
    rule_free(&R);
    return -1;
  }
  double lw[ORC_MAXDIM], hg[ORC_MAXDIM];
  for (int j = 0; j < d; ++j) {
    lw[j] = 0.;
    hg[j] = 1.;
  }
  size_t cap = maxeval / (size_t)R.F + 4;
  creg_t* heap = (creg_t*)malloc(sizeof(creg_t) * cap);
  size_t nh = 0;
  creg_t r0;
  for (int j = 0; j < d; ++j) {
    r0.lo[j] = 0.;
    r0.len[j] = 1.;
  }
  eval_one(&R, fam, prm, lw, hg, r0.lo, r0.len, &r0.avg, &r0.err, &r0.bisect,
           NULL, NULL);
  heap_push(heap, &nh, &r0);
  double tot_avg = r0.avg, tot_err = r0.err;
  size_t neval = (size_t)R.F, nsplit = 0;
  int status = 1;
  for (;;) {
    if (tot_err <= fmax(epsrel * fabs(tot_avg), epsabs)) {
      status = 0;
      break;
    }
    if (neval + 2 * (size_t)R.F > maxeval || nh + 2 > cap) break;
    creg_t par = heap_pop(heap, &nh);
    creg_t ch[2];
    for (int c = 0; c < 2; ++c) {
      ch[c] = par;
      const int b = par.bisect;
      ch[c].len[b] = par.len[b] / 2;
      ch[c].lo[b] = par.lo[b] + c * ch[c].len[b];
      eval_one(&R, fam, prm, lw, hg, ch[c].lo, ch[c].len, &ch[c].avg,
               &ch[c].err, &ch[c].bisect, NULL, NULL);
    }
    /* two-level refinement of the children (same update as RefineError) */
    const double diff = fabs(.25 * (ch[0].avg + ch[1].avg - par.avg));
    const double E = ch[0].err + ch[1].err;
    if (E > 0) {
      const double c = 1 + 2 * diff / E;
      ch[0].err *= c;
      ch[1].err *= c;
    }
    ch[0].err += diff;
    ch[1].err += diff;
    tot_avg += ch[0].avg + ch[1].avg - par.avg;
    tot_err += ch[0].err + ch[1].err - par.err;
    heap_push(heap, &nh, &ch[0]);
    heap_push(heap, &nh, &ch[1]);
    neval += 2 * (size_t)R.F;
    nsplit++;
  }
  memset(out, 0, sizeof(*out));
  out->estimate = tot_avg;
  out->errorest = tot_err;
  out->neval = neval;
  out->nregions = 1 + 2 * nsplit;
  out->status = status;
  out->iters = nsplit;
  out->lastPhase = -1;
  free(heap);
  rule_free(&R);
  return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
