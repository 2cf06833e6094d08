// Checks the hand-derived VJPs of diablo3_test_amd/csrc/les_point.h against
// central finite differences of the point models on random inputs:
// <ybar, (f(x + e v) - f(x - e v)) / 2e>  vs  <vjp(ybar), v>.
// Prints one line per case: "<model> <type> <max relative error>".
#include <cmath>
#include <cstdio>
#include <random>

#include "../diablo3_test_amd/csrc/les_point.h"

using namespace lespt;

static std::mt19937_64 rng(12345);
static double rnd(double a = 1.0) { return std::uniform_real_distribution<double>(-a, a)(rng); }

template <typename T>
static double* as(T& x) { return reinterpret_cast<double*>(&x); }

int main() {
  const double eps = 1e-6;
  for (int type : {1, 4, 5}) {
    double worst = 0.0;
    int trials = 0;
    for (int t = 0; t < 400; ++t) {
      VelPlane<double> P;
      P.type = type;
      P.gyf_in = (t % 4) != 1;
      P.gy_in = (t % 4) != 2;
      P.nt_in = P.gy_in && (t % 5) != 0;
      P.cr_in = (t % 7) != 0;
      P.dyj = 0.05 + 0.1 * std::fabs(rnd());
      P.dyfj = 0.05 + 0.1 * std::fabs(rnd());
      P.dyfm = 0.05 + 0.1 * std::fabs(rnd());
      P.dysu = 2 * (0.05 + 0.1 * std::fabs(rnd()));
      P.dxs = 0.3 + 0.2 * std::fabs(rnd());
      P.dzs = 0.2 + 0.2 * std::fabs(rnd());
      P.dyf_fac = -0.1 - 0.1 * std::fabs(rnd());
      P.dy_fac = -0.1 - 0.1 * std::fabs(rnd());
      VelIn<double> x, v, xb = {};
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) {
        as(x)[i] = rnd();
        as(v)[i] = rnd();
      }
      VelOut<double> yb;
      for (int i = 0; i < 8; ++i) as(yb)[i] = rnd();
      VelIn<double> xp = x, xm = x;
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) {
        as(xp)[i] += eps * as(v)[i];
        as(xm)[i] -= eps * as(v)[i];
      }
      VelOut<double> yp, ym;
      vel_point(P, xp, yp);
      vel_point(P, xm, ym);
      // skip points whose AMD clip changes inside the difference stencil
      VelOut<double> y0;
      vel_point(P, x, y0);
      double fd = 0.0;
      for (int i = 0; i < 8; ++i) fd += as(yb)[i] * (as(yp)[i] - as(ym)[i]) / (2 * eps);
      vel_point_vjp(P, x, yb, xb);
      double ad = 0.0;
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) ad += as(xb)[i] * as(v)[i];
      const bool kink = (type != 1) && ((vel_gyf_scalar(P, xp) == 0.0) != (vel_gyf_scalar(P, xm) == 0.0) ||
                                        (vel_gy_scalar(P, xp) == 0.0) != (vel_gy_scalar(P, xm) == 0.0));
      if (kink) continue;
      const double err = std::fabs(fd - ad) / (std::fabs(fd) + std::fabs(ad) + 1e-12);
      if (err > worst) worst = err;
      ++trials;
    }
    std::printf("vel %d %.3e %d\n", type, worst, trials);
  }
  {
    double worst = 0.0;
    int trials = 0;
    for (int t = 0; t < 400; ++t) {
      ThPlane<double> P;
      P.gyf_in = (t % 3) != 1;
      P.gy_in = (t % 3) != 2;
      P.dxs = 0.3 + 0.2 * std::fabs(rnd());
      P.dzs = 0.2 + 0.2 * std::fabs(rnd());
      P.dyj = 0.05 + 0.1 * std::fabs(rnd());
      P.dyfj = 0.05 + 0.1 * std::fabs(rnd());
      P.dyfm = 0.05 + 0.1 * std::fabs(rnd());
      P.dyf_fac = -0.1 - 0.1 * std::fabs(rnd());
      P.dy_fac = -0.1 - 0.1 * std::fabs(rnd());
      ThIn<double> x, v, xb = {};
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) {
        as(x)[i] = rnd();
        as(v)[i] = rnd();
      }
      ThOut<double> yb = {rnd(), rnd(), rnd()};
      ThIn<double> xp = x, xm = x;
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) {
        as(xp)[i] += eps * as(v)[i];
        as(xm)[i] -= eps * as(v)[i];
      }
      ThOut<double> yp, ym, y0, y1;
      th_point(P, xp, yp);
      th_point(P, xm, ym);
      // kink detection: a clipped output on one side only
      ThPlane<double> Pa = P;
      th_point(Pa, x, y0);
      bool kink = false;
      for (int i = 0; i < 3; ++i)
        if ((as(yp)[i] == 0.0) != (as(ym)[i] == 0.0)) kink = true;
      if (kink) continue;
      (void)y1;
      double fd = 0.0;
      for (int i = 0; i < 3; ++i) fd += as(yb)[i] * (as(yp)[i] - as(ym)[i]) / (2 * eps);
      th_point_vjp(P, x, yb, xb);
      double ad = 0.0;
      for (int i = 0; i < (int)(sizeof(x) / sizeof(double)); ++i) ad += as(xb)[i] * as(v)[i];
      const double err = std::fabs(fd - ad) / (std::fabs(fd) + std::fabs(ad) + 1e-12);
      if (err > worst) worst = err;
      ++trials;
    }
    std::printf("th 5 %.3e %d\n", worst, trials);
  }
  return 0;
}
