// rule.cpp -- degree-9 rule tables, see rule.hpp.
#include "rule.hpp"

#include <cmath>

namespace pagani {

namespace {

// Generator magnitudes lambda_1..lambda_5, Rule.cuh:201-205.
constexpr double kLam[5] = {.4779536579022695061928604197171830064732,
                            .2030285873691198677998034402373279133258,
                            .4476273546261781288207704806530998539285, .125,
                            .3430378987808781457001426145164678603407};

// Weights as polynomials in ndim, Rule.cuh:36-147.  Every expression keeps
// the reference's operation order, so the doubles are identical.
void weights(int d, double (&w)[PAGANI_NSETS][PAGANI_NRULES])
{
  const double D = static_cast<double>(d);
  // centre (set 0)
  const double cubic_a = D * (D * (D * -.002361170967785511788400941242259231309691 +
                                   .1141539002385732526821323741697655347686) +
                              -.6383392007670238909386026193674701393074);
  const double cubic_b = D * (D * (D * -.001432401703339912514196154599769007103671 +
                                   .05747150786448972594860897296200006759892) +
                              -.1422510457143424323449521620935950679394);
  const double a0 = .7484998850468520800423030047583803945205;
  const double b0 = -.06287502873828697998942424881040490136987;
  w[0][0] = cubic_a + a0;
  w[0][1] = cubic_b - b0;
  w[0][2] = D * .2545911332489590890011611142429070613156 - (cubic_b - b0);
  w[0][3] = ((D * (D * -1.207328566678236261002219995185143356737 +
                   .8956736576416067650809467826488567200939) -
              1.0) +
             cubic_a) +
            a0;
  w[0][4] = (D * -.3647935698604914666100134551377381205297 + 1.0) - (cubic_a + a0);
  // +-lambda1 on one axis (set 1)
  const double quad_g = D * (D * .003541756451678267682601411863388846964536 +
                             -.07260936739589367960492815865074633743652);
  const double quad_h = D * (D * .002148602555009868771294231899653510655506 +
                             -.03226856389295394999786630399875134318006);
  const double g0 = .1055749162521899101218622863269817454540;
  const double h0 = .01063678399023121748083624225818915724455;
  w[1][0] = quad_g + g0;
  w[1][1] = quad_h + h0;
  w[1][2] = .01468910249614349017540783437728097691502 - (quad_h + h0);
  w[1][3] = ((D * .5113470834646759143109387357149329909126 +
              .4597644812080634464633352781605214342691) +
             quad_g) +
            g0;
  w[1][4] = .1823967849302457333050067275688690602649 - (quad_g + g0);
  // +-lambda2 on one axis (set 2)
  const double lin_p = D * -.04508628929435784075980562738240804429658 +
                       .2141588352435279340097929526588394300172;
  const double lin_q = D * -.02735154652654564472203690086290223507436 +
                       .05494106704871123410060080562462135546101;
  w[2][0] = lin_p;
  w[2][1] = lin_q;
  w[2][2] = .1193759620257077529708962121565290178730 - lin_q;
  w[2][3] = D * .6508951939192025059314756320878023215278 +
            .1474493982943446016775696826942585013243;
  w[2][4] = -lin_p;
  // +-lambda3 on one axis (set 3)
  w[3][0] = .05769338449097348357291272840392627722165;
  w[3][1] = .03499962660214358382244159694487155861542;
  w[3][2] = -.05769338449097348357291272840392627722165;
  w[3][3] = -1.386862771927828143599782668709014266770;
  w[3][4] = -.05769338449097348357291272840392627722165;
  // +-lambda4 on one axis (set 4): only null rule 2 uses it
  w[4][0] = 0.;
  w[4][1] = 0.;
  w[4][2] = -.2386668732575008878964134721962088068396;
  w[4][3] = 0.;
  w[4][4] = 0.;
  // (+-lambda1, +-lambda1) (set 5)
  const double t0 = .01553241727660705326386197156586357005224;
  const double t1 = .003532809960709087023561817517751309380604;
  w[5][0] = t0 - D * .003541756451678267682601411863388846964536;
  w[5][1] = t1 - D * .002148602555009868771294231899653510655506;
  w[5][2] = -(t1 - D * .002148602555009868771294231899653510655506);
  w[5][3] = (.09231719987444221619017126187763868745587 + t0) -
            D * .003541756451678267682601411863388846964536;
  w[5][4] = -(t0 - D * .003541756451678267682601411863388846964536);
  // (+-lambda1, +-lambda2) (set 6)
  w[6][0] = .02254314464717892037990281369120402214829;
  w[6][1] = .01367577326327282236101845043145111753718;
  w[6][2] = -.01367577326327282236101845043145111753718;
  w[6][3] = -.3254475969596012529657378160439011607639;
  w[6][4] = -.02254314464717892037990281369120402214829;
  // (+-lambda1, +-lambda1, +-lambda1) (set 7)
  w[7][0] = .001770878225839133841300705931694423482268;
  w[7][1] = .001074301277504934385647115949826755327753;
  w[7][2] = -.001074301277504934385647115949826755327753;
  w[7][3] = .001770878225839133841300705931694423482268;
  w[7][4] = -.001770878225839133841300705931694423482268;
  // all coordinates +-lambda5 (set 8)
  const double corners = static_cast<double>(1 << d);
  const double u0 = .2515001149531479199576969952416196054795;
  const double v0 = -.06287502873828697998942424881040490136987;
  w[8][0] = u0 / corners;
  w[8][1] = v0 / corners;
  w[8][2] = -(v0 / corners);
  w[8][3] = u0 / corners;
  w[8][4] = -(u0 / corners);
}

}  // namespace

int feval_points(int d)
{
  if (d < PAGANI_MIN_DIM || d > PAGANI_MAX_DIM) return -1;
  return 1 + 8 * d + 2 * d * (d - 1) + 4 * d * (d - 1) + 4 * d * (d - 1) * (d - 2) / 3 +
         (1 << d);
}

bool build_rule(int d, RuleTables& R)
{
  const int F = feval_points(d);
  if (F < 0) return false;
  R.ndim = d;
  R.feval = F;
  weights(d, R.w);
  for (int i = 0; i < 5; ++i) R.lambda[i] = kLam[i];
  const int cnt[PAGANI_NSETS] = {1,
                                 2 * d,
                                 2 * d,
                                 2 * d,
                                 2 * d,
                                 2 * d * (d - 1),
                                 4 * d * (d - 1),
                                 4 * d * (d - 1) * (d - 2) / 3,
                                 1 << d};
  int start = 0;
  for (int s = 0; s < PAGANI_NSETS; ++s) {
    R.count[s] = cnt[s];
    R.set_begin[s] = start;
    start += cnt[s];
  }
  R.set_begin[PAGANI_NSETS] = start;

  // null-rule scale and norm (Rule.cuh:256-269); columns 1..3 are used
  for (int s = 0; s < PAGANI_NSETS; ++s) {
    for (int r = 1; r < PAGANI_NRULES - 1; ++r) {
      const double sc = (R.w[s][r] == 0) ? 100 : -R.w[s][r + 1] / R.w[s][r];
      double sum = 0;
      for (int x = 0; x < PAGANI_NSETS; ++x)
        sum += cnt[x] * std::fabs(R.w[x][r + 1] + sc * R.w[x][r]);
      R.scale[s][r] = sc;
      R.norm[s][r] = 1 / sum;
    }
  }

  // Generator table for sets 3..8 (every signed permutation of the orbit
  // seed, Rule.cuh:217-249, 438-505).
  R.gtab.clear();
  auto point = [&](auto&& fill) {
    const std::size_t base = R.gtab.size();
    R.gtab.resize(base + d, 0.0);
    fill(&R.gtab[base]);
  };
  for (int s = 3; s <= 4; ++s)
    for (int j = 0; j < d; ++j)
      for (int sg = 0; sg < 2; ++sg)
        point([&](double* g) { g[j] = sg ? -kLam[s - 1] : kLam[s - 1]; });
  for (int j = 0; j < d; ++j)
    for (int k = j + 1; k < d; ++k)
      for (int sg = 0; sg < 4; ++sg)
        point([&](double* g) {
          g[j] = (sg & 1) ? -kLam[0] : kLam[0];
          g[k] = (sg & 2) ? -kLam[0] : kLam[0];
        });
  for (int j = 0; j < d; ++j)
    for (int k = 0; k < d; ++k) {
      if (k == j) continue;
      for (int sg = 0; sg < 4; ++sg)
        point([&](double* g) {
          g[j] = (sg & 1) ? -kLam[0] : kLam[0];
          g[k] = (sg & 2) ? -kLam[1] : kLam[1];
        });
    }
  for (int j = 0; j < d; ++j)
    for (int k = j + 1; k < d; ++k)
      for (int l = k + 1; l < d; ++l)
        for (int sg = 0; sg < 8; ++sg)
          point([&](double* g) {
            g[j] = (sg & 1) ? -kLam[0] : kLam[0];
            g[k] = (sg & 2) ? -kLam[0] : kLam[0];
            g[l] = (sg & 4) ? -kLam[0] : kLam[0];
          });
  for (int c = 0; c < (1 << d); ++c)
    point([&](double* g) {
      for (int j = 0; j < d; ++j) g[j] = ((c >> j) & 1) ? -kLam[4] : kLam[4];
    });
  return static_cast<int>(R.gtab.size()) == (F - R.set_begin[3]) * d;
}

double point_generator(const RuleTables& R, int p, int axis)
{
  if (p == 0) return 0.0;
  if (p < R.set_begin[3]) {
    const bool two = p >= R.set_begin[2];
    const int k = p - (two ? R.set_begin[2] : R.set_begin[1]);
    if (k / 2 != axis) return 0.0;
    const double lam = two ? R.lambda[1] : R.lambda[0];
    return (k & 1) ? -lam : lam;
  }
  return R.gtab[static_cast<std::size_t>(p - R.set_begin[3]) * R.ndim + axis];
}

}  // namespace pagani
