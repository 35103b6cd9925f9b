// Host-side tables of libmfbem: Gauss-Legendre rules, quasi-singular estimator parameters and the
// complex coefficients of the harmonic fundamental solution.  Computed once per call on the host
// and handed to the kernels as plain data.
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mfbem_types.h"

namespace mfb {

// Gauss-Legendre nodes and weights by Newton iteration on P_n in binary128, rounded to 17
// significant digits: the reference holds its rules as 17-digit data statements
// (lib/fbem/src/resources_quad_rules/gl11.rc, gl01.rc), and rounding through the same decimal
// precision reproduces those doubles.
static double round17(long double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.16Le", v);
  return std::strtod(buf, nullptr);
}

void host_gl_tables(double* x11, double* w11, double* x01, double* w01) {
  typedef __float128 q_t;
  std::memset(x11, 0, sizeof(double) * 33 * 32);
  std::memset(w11, 0, sizeof(double) * 33 * 32);
  std::memset(x01, 0, sizeof(double) * 33 * 32);
  std::memset(w01, 0, sizeof(double) * 33 * 32);
  for (int n = 1; n <= 32; n++) {
    long double xs[32], ws[32];
    for (int i = 0; i < (n + 1) / 2; i++) {
      q_t z = (q_t)std::cos(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
      q_t p0 = 1, p1 = z, dp = 1;
      for (int it = 0; it < 200; it++) {
        p0 = 1;
        p1 = z;
        for (int k = 1; k < n; k++) {
          q_t p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
          p0 = p1;
          p1 = p2;
        }
        dp = (n == 1) ? (q_t)1 : n * (z * p1 - p0) / (z * z - 1);
        q_t dz = p1 / dp;
        z -= dz;
        if (dz < 0) dz = -dz;
        if (dz < (q_t)1e-32L) break;
      }
      p0 = 1;
      p1 = z;
      for (int k = 1; k < n; k++) {
        q_t p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
      }
      dp = (n == 1) ? (q_t)1 : n * (z * p1 - p0) / (z * z - 1);
      q_t w = 2 / ((1 - z * z) * dp * dp);
      xs[i] = (long double)(-z);
      xs[n - 1 - i] = (long double)z;
      ws[i] = ws[n - 1 - i] = (long double)w;
    }
    if (n % 2 == 1) xs[n / 2] = 0;
    for (int i = 0; i < n; i++) {
      x11[n * 32 + i] = round17(xs[i]);
      w11[n * 32 + i] = round17(ws[i]);
      x01[n * 32 + i] = round17(0.5L * (1 + xs[i]));
      w01[n * 32 + i] = round17(0.5L * ws[i]);
    }
  }
}

// Fitted N(d) curves of the reference, [curve][coefficient]
// (lib/fbem/src/quasisingular_integration.f90:181-292).  These published fit coefficients define
// which quadrature order the reference picks, so they have to be the same numbers.
static const double kLnrA0[3][3] = {{2.511371433253553e-1, -8.318131267946077e-2, -1.973643360180158e-3},
                                    {1.588915580092415e-1, -9.469173675969373e-2, -2.500979779162619e-3},
                                    {1.303338290205561e-1, -9.041069352366727e-2, -2.302123487272636e-3}};
static const double kLnrA1[3][3] = {{-8.301769338383449e-1, -7.458331284021328e-2, -1.992535165956408e-3},
                                    {-8.575272011921337e-1, -6.866562987114574e-2, -1.735469217262563e-3},
                                    {-5.785210418300009e-1, -3.107089548772586e-2, -4.204818399993008e-4}};
static const double kLnrB1[3][3] = {{-1.114783338380396e-1, -6.169581618278328e-2, -1.873698813959166e-3},
                                    {-3.918234478283774e-1, -6.478096580939043e-2, -1.657459305149584e-3},
                                    {-1.933844506192423e-1, -3.303326826061181e-2, -6.497990057394490e-4}};
static const double kRnA0[3][6] = {
    {2.443943976922257e-1, -7.845125775746743e-2, 3.895326800087784e-2, 8.783124822346978e-4, -1.719172638314362e-3, -1.575912106080523e-3},
    {1.855869272054077e-1, -8.441784317192081e-2, 3.992519918023225e-2, 1.003034542932314e-3, -2.012460287158036e-3, -1.596684136739970e-3},
    {1.964994120809251e-1, -7.533682025056084e-2, 6.300967117113785e-2, 1.453521059340246e-3, -1.653933461489553e-3, -2.613368441217468e-3}};
static const double kRnA1[3][6] = {
    {-8.421216556191472e-1, -8.778262996280104e-2, 1.679496012553309e-2, 6.522695414940627e-4, -2.649807213627071e-3, -8.472009602597869e-4},
    {-8.012899104725332e-1, -7.060561516342915e-2, 5.713299243166406e-3, -1.952251227204736e-4, -1.984455335909311e-3, -6.680823196695776e-4},
    {-5.891778953073477e-1, -4.515354064860566e-2, 5.728946548799128e-3, 1.106131193536092e-6, -1.179100966502172e-3, -5.355782068751523e-4}};
static const double kRnB1[3][6] = {
    {-2.693032219216048e-2, -6.138311391493901e-2, 2.273921394673689e-2, 9.789745450155518e-4, -2.030377872751224e-3, -9.059289455187508e-4},
    {-2.353515879579821e-1, -5.141856801187955e-2, 1.693100431819354e-2, -1.491381226735667e-5, -1.330369292483563e-3, -1.264704521827383e-3},
    {-9.828561879516765e-2, -3.074365737570111e-2, 1.306905121742232e-2, 6.065437465834727e-5, -7.853234287066991e-4, -1.025889829482388e-3}};
static const double kTLnrA0[3][3] = {{2.5834148732821033e-01, -8.8865167808209675e-02, -2.3640194116454241e-03},
                                     {5.1283007739426911e-01, -1.1479138136155583e-02, 1.7167160355773211e-03},
                                     {4.1061632629507161e-01, -3.0669024691563598e-02, 5.1920314153891956e-04}};
static const double kTLnrA1[3][3] = {{-9.1721308433764848e-01, -6.2837979237008845e-02, -1.1580362636310136e-03},
                                     {-1.0632414684959128e+00, -1.1633428696469847e-01, -4.4173578665990879e-03},
                                     {-1.1740312565911180e+00, -1.3946434974037192e-01, -5.4264705520737193e-03}};
static const double kTLnrB1[3][3] = {{-4.9431525769444484e-01, -6.5628053690049357e-02, -1.2683068366630545e-03},
                                     {-9.1742055835524161e-01, -1.4448614058834844e-01, -5.5849846784841173e-03},
                                     {-1.1668939380105636e+00, -1.7686541251294283e-01, -6.9562836980345301e-03}};
static const double kTRnA0[3][6] = {
    {2.9379042486857926e-01, -7.6518438657183094e-02, 3.0011332045426935e-02, 5.8900790667561719e-04, -1.7642732320416377e-03, -1.3634192833061319e-03},
    {1.1155167367380846e-01, -9.0043220465500617e-02, 3.9605615471752335e-02, 1.0669688965738889e-03, -2.2428131063082419e-03, -1.6030309471606837e-03},
    {9.5587534756324991e-02, -8.5841370957250140e-02, 5.9724590614136340e-02, 1.4101322853742803e-03, -2.0424217927592597e-03, -2.4264102214015559e-03}};
static const double kTRnA1[3][6] = {
    {-9.8339088502232042e-01, -8.6761356304300027e-02, 2.6028760963988956e-02, 7.7677507397256337e-04, -2.3619290417702245e-03, -1.2891431252005461e-03},
    {-6.5856321275500784e-01, -4.7354561095619759e-02, 1.9423558756522371e-03, -2.9087437598165618e-04, -1.1859755938049318e-03, -4.5988734937654318e-04},
    {-2.4634605754204883e-01, 1.0051411705899942e-02, -1.1994978255365170e-02, -1.0022065157916291e-03, 7.9639034155470554e-04, -1.8103634920482468e-04}};
static const double kTRnB1[3][6] = {
    {-4.5453529921363467e-01, -7.3941529478921839e-02, 3.7672240427426076e-02, 9.8276802658567980e-04, -1.8897770518175436e-03, -1.8079562352345093e-03},
    {-2.6610069546799336e-01, -3.6048387668689889e-02, 1.3288129149913631e-02, -1.8099231920766130e-04, -7.6015457149682889e-04, -1.2247286683263207e-03},
    {7.5605400079843565e-02, 2.4021121564322980e-02, -9.3693044556669332e-03, -1.3532246088221670e-03, 1.3564674989692248e-03, -7.4467072775070487e-04}};

static double quad1(const double* c, double x) { return c[0] + c[1] * x + c[2] * (x * x); }
static double quad2(const double* c, double x, double y) {
  return c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * (x * x) + c[5] * (y * y);
}

// fbem_qs_calculate_parameters (quasisingular_integration.f90:318-400).  log10(30.) and log10(2.)
// are default-kind (single precision) literals there; they are rounded to float here as well.
void host_qsi_params(double relative_error, QsiParams* p) {
  const double l30 = (double)(float)1.4771212547196624, l2 = (double)(float)0.3010299956639812;
  double l = std::log10(std::fabs(relative_error));
  if (l > -3.0) l = -3.0;
  if (l < -15.0) l = -15.0;
  for (int c = 0; c < 3; c++) {
    p->a0[c][0] = quad1(kLnrA0[c], l);
    p->a1[c][0] = quad1(kLnrA1[c], l);
    p->b1[c][0] = quad1(kLnrB1[c], l);
    p->ta0[c][0] = quad1(kTLnrA0[c], l);
    p->ta1[c][0] = quad1(kTLnrA1[c], l);
    p->tb1[c][0] = quad1(kTLnrB1[c], l);
    for (int i = 1; i <= 7; i++) {
      p->a0[c][i] = quad2(kRnA0[c], l, (double)i);
      p->a1[c][i] = quad2(kRnA1[c], l, (double)i);
      p->b1[c][i] = quad2(kRnB1[c], l, (double)i);
      p->ta0[c][i] = quad2(kTRnA0[c], l, (double)i);
      p->ta1[c][i] = quad2(kTRnA1[c], l, (double)i);
      p->tb1[c][i] = quad2(kTRnB1[c], l, (double)i);
    }
    for (int i = 0; i <= 7; i++) {
      p->dmin[c][i] = std::pow(10.0, (l30 - p->a0[c][i]) / (p->a1[c][i] - l30 * p->b1[c][i]));
      p->dmax[c][i] = std::pow(10.0, (l2 - p->a0[c][i]) / (p->a1[c][i] - l2 * p->b1[c][i]));
      p->tdmin[c][i] = std::pow(10.0, (l30 - p->ta0[c][i]) / (p->ta1[c][i] - l30 * p->tb1[c][i]));
      p->tdmax[c][i] = std::pow(10.0, (l2 - p->ta0[c][i]) / (p->ta1[c][i] - l2 * p->tb1[c][i]));
      if (p->tdmin[c][i] < 1e-6) p->tdmin[c][i] = 1e-6;
      if (p->tdmin[c][i] > p->tdmax[c][i]) p->tdmin[c][i] = 1e-6;
    }
  }
}

// fbem_bem_harela3d_calculate_parameters (bem_harela3d.f90:143-287), the u* and t* coefficients.
void host_har_params(double lam_re, double lam_im, double mu_re, double mu_im, double rho,
                     double omega, HarParams* p) {
  typedef std::complex<double> z;
  const z I(0.0, 1.0);
  const double one_4pi = 0.07957747154594767280411105048;
  z lambda(lam_re, lam_im), mu(mu_re, mu_im);
  z c1 = std::sqrt((lambda + 2.0 * mu) / rho), c2 = std::sqrt(mu / rho);
  z k1 = omega / c1, k2 = omega / c2;
  z r = (c2 * c2) / (c1 * c1), ik1 = I * k1, ik2 = I * k2, c1s = c1 * c1, c2s = c2 * c2;
  z psi[6] = {0.5 * (1.0 + r), -1.0 / 3.0 * (2.0 / c2 + c2s / (c1s * c1)) * I * omega,
              I * k1 / (k2 * k2), 1.0 / ik2, 1.0 / (k2 * k2), -1.0 / (k2 * k2)};
  z chi[6] = {-0.5 * (1.0 - r), -r, -3.0 * r / ik1, 3.0 / ik2, -3.0 * r / (ik1 * ik1),
              3.0 / (ik2 * ik2)};
  z T1[10] = {3.0 * (r - 1.0), -0.25 * (1.0 / c2s - c2s / (c1s * c1s)) * (omega * omega),
              -2.0 * I * k1 * r, 2.0 * I * k2, -12.0 * r, 12.0, -r * 30.0 / ik1, 30.0 / ik2,
              -r * 30.0 / (ik1 * ik1), 30.0 / (ik2 * ik2)};
  z T2[9] = {-r, -0.25 * (1.0 / c2s + c2s / (c1s * c1s)) * (omega * omega), -I * k2, 2.0 * r,
             -3.0, r * 6.0 / ik1, -6.0 / ik2, r * 6.0 / (ik1 * ik1), -6.0 / (ik2 * ik2)};
  z T3[9] = {r, 0.25 * (3.0 * c2s / (c1s * c1s) - 2.0 / c1s + 1.0 / c2s) * (omega * omega),
             (2.0 * r - 1.0) * I * k1, 4.0 * r - 1.0, -2.0, r * 6.0 / ik1, -6.0 / ik2,
             r * 6.0 / (ik1 * ik1), -6.0 / (ik2 * ik2)};
  auto put = [](double* d, z v) { d[0] = v.real(); d[1] = v.imag(); };
  put(p->k1, k1);
  put(p->k2, k2);
  for (int i = 0; i < 6; i++) { put(p->psi[i], psi[i]); put(p->chi[i], chi[i]); }
  for (int i = 0; i < 10; i++) put(p->T1[i], T1[i]);
  for (int i = 0; i < 9; i++) { put(p->T2[i], T2[i]); put(p->T3[i], T3[i]); }
  put(p->cte_u, one_4pi / mu);
  put(p->cte_t, z(one_4pi, 0.0));
}

}  // namespace mfb
