// Host-side k-space kernel generators: fvlj4 (fforma.f90), wcgk (initcg.f90),
// kalfs (DFT4He3d.f90:566-570) and Vq by the spline Fourier transform of
// FT_V_spline.f90 applied to a cored Morse potential.
#pragma once
#include <cmath>
#include <complex>
#include <thread>
#include <vector>

namespace hetd {
namespace kinit {

constexpr double kPi = 3.141592653589793238462643383279502884;

// FFTW-ordered frequency with +pmax at n/2 (DFT4He3d.f90:321-339).
inline double pfreq(int i, int n, double h) {
  const double hp = 1.0 / (n * h);
  const double pmax = 1.0 / (2.0 * h);
  return i <= n / 2 ? hp * i : -pmax + hp * (i - n / 2);
}

// Sine integral Si(x), x >= 0.  Power series for x <= 4; above, the
// continued fraction of E1(ix) = e^{-ix}/(1+ix- 1/(3+ix- 4/(5+ix- ...)))
// evaluated with the modified Lentz method, Si = pi/2 + Im(E1(ix)).
inline double sine_integral(double x) {
  if (x < 0) return -sine_integral(-x);
  if (x <= 4.0) {
    double term = x, sum = x;
    const double x2 = x * x;
    for (int k = 1; k < 60; ++k) {
      term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
      const double add = term / (2.0 * k + 1.0);
      sum += add;
      if (std::fabs(add) < 1e-18 * std::fabs(sum)) break;
    }
    return sum;
  }
  using cd = std::complex<double>;
  const double tiny = 1e-300;
  cd b(1.0, x);
  cd c(1.0 / tiny, 0.0);
  cd d = 1.0 / b;
  cd h = d;
  for (int i = 2; i < 100000; ++i) {
    const double a = -double(i - 1) * double(i - 1);
    b += 2.0;
    d = 1.0 / (a * d + b);
    c = b + a / c;
    const cd del = c * d;
    h *= del;
    if (std::abs(del.real() - 1.0) + std::abs(del.imag()) < 1e-17) break;
  }
  h *= cd(std::cos(x), -std::sin(x));
  return kPi * 0.5 + h.imag();
}

inline double sinint(double a, double x) { return kPi * 0.5 - sine_integral(a * x); }

// int_x^inf sin(a r)/r^11 dr (fforma.f90:73-100)
inline double sinx11(double a, double x) {
  const double y = a * x, y2 = y * y;
  const double db = 40320.0 - 720.0 * y2 + 24.0 * (y2 * y2) - 2.0 * std::pow(y2, 3) + 1.0 * std::pow(y2, 4);
  const double dc = 362880.0 - 5040.0 * y2 + 120.0 * (y2 * y2) - 6.0 * std::pow(y2, 3) + 1.0 * std::pow(y2, 4);
  const double x9 = std::pow(x, 9);
  const double res1 = (a * db * std::cos(y) + dc * std::sin(y) / x) / (3628800.0 * x9);
  const double res2 = -(std::pow(a, 10) * sinint(a, x)) / 3628800.0;
  return res1 + res2;
}

// int_x^inf sin(a r)/r^5 dr (fforma.f90:106-133)
inline double sinx5(double a, double x) {
  const double y = a * x, y2 = y * y;
  const double db = -2.0 + 1.0 * y2, dc = -6.0 + 1.0 * y2;
  const double res1 = -(a * db * std::cos(y) + dc * std::sin(y) / x) / (24.0 * std::pow(x, 3));
  const double res2 = (std::pow(a, 4) * sinint(a, x)) / 24.0;
  return res1 + res2;
}

// Maron cubic spline in global-polynomial form (mates.f:36-116).
inline void spline(const std::vector<double>& x, const std::vector<double>& y, double s1, double sn,
                   std::vector<double>& A, std::vector<double>& B, std::vector<double>& C,
                   std::vector<double>& D) {
  const int n = (int)x.size(), n1 = n - 1, n2 = n - 2;
  std::vector<double> a(n, 0.0), b(n, 0.0), d(n, 0.0);
  for (int i = 0; i < n1; ++i) {
    a[i] = x[i + 1] - x[i];
    d[i] = (y[i + 1] - y[i]) / a[i];
  }
  for (int i = 1; i <= n2; ++i) {
    b[i - 1] = 2.0 * (a[i - 1] + a[i]);
    const int k = n1 - i + 1;
    d[k - 1] = 6.0 * (d[k - 1] - d[k - 2]);
  }
  d[1] = d[1] - a[0] * s1;
  d[n1 - 1] = d[n1 - 1] - a[n1 - 1] * sn;
  for (int i = 2; i <= n2; ++i) {
    const double r = a[i - 1] / b[i - 2];
    b[i - 1] = b[i - 1] - r * a[i - 1];
    d[i] = d[i] - r * d[i - 1];
  }
  d[n1 - 1] = d[n1 - 1] / b[n2 - 1];
  for (int i = 2; i <= n2; ++i) {
    const int k = n1 - i + 1;
    d[k - 1] = (d[k - 1] - a[k - 1] * d[k]) / b[k - 2];
  }
  d[n - 1] = sn;
  A.assign(n, 0.0);
  B.assign(n, 0.0);
  C.assign(n, 0.0);
  D.assign(n, 0.0);
  double si1 = s1;
  for (int i = 0; i < n1; ++i) {
    const double si = si1;
    si1 = d[i + 1];
    const double h = a[i], hi = 1.0 / h;
    const double xi = x[i], xi1 = x[i + 1];
    A[i] = (hi / 6.0) * (si * std::pow(xi1, 3) - si1 * std::pow(xi, 3)) + hi * (y[i] * xi1 - y[i + 1] * xi) +
           (h / 6.0) * (si1 * xi - si * xi1);
    B[i] = (hi / 2.0) * (si1 * (xi * xi) - si * (xi1 * xi1)) + hi * (y[i + 1] - y[i]) + (h / 6.0) * (si - si1);
    C[i] = (hi / 2.0) * (si * xi1 - si1 * xi);
    D[i] = (hi / 6.0) * (si1 - si);
  }
}

// Bisection interval search (mates.f:4-34); returns a 0-based index.
inline int findi(const std::vector<double>& xs, double xc) {
  const int n = (int)xs.size();
  if (xc > xs[n - 1]) return n - 2;
  if (xc < xs[0]) return 0;
  int i = 1, i1 = n;
  for (;;) {
    const int it = (i + i1) / 2;
    if (xc > xs[it - 1]) i = it;
    if (xc <= xs[it - 1]) i1 = it;
    if (i1 - i <= 1) break;
  }
  return i - 1;
}

struct MorsePot {
  double D, re, alpha, rc, umax;
  double operator()(double r) const {
    if (r < rc) return umax;
    const double e = std::exp(-alpha * (r - re));
    return D * (e * e - 2.0 * e);
  }
};

// Spline Fourier transform of a radial potential (FT_V_spline.f90:14-276), in
// three phases so that the momentum sums can also run on the device:
//   FtvSetup   the cubic spline of V on 100001 points in [rc + 1e-8, 2000 A]
//              (FT_V_spline.f90:84-123) and the momentum grid P;
//   ftv_sums   V(p) = 4 pi/p [vp(p) + sum over spline pieces of the piece's
//              analytic Fourier integral] for every p of the grid (:151-225);
//   FtvSpline  the small-q redefinition (:230-264) and the spline over p used
//              for lookups (:47-53, 268).
struct FtvSetup {
  static constexpr int N = 100001;
  std::vector<double> x, a, b, c, d, P;
  double rc = 0.0, fc = 0.0;

  FtvSetup(const MorsePot& pot, double pmax, double hp) {
    const double rmax = 2000.0, eps = 1e-8, dr = 0.5;
    rc = pot.rc;
    fc = pot.umax;
    const double X1 = rc + eps, XN = rmax;
    const double S1 = (pot(X1) - 2.0 * pot(X1 + dr) + pot(X1 + dr + dr)) / (dr * dr);
    const double Sn = (pot(XN - dr) - 2.0 * pot(XN) + pot(XN + dr)) / (dr * dr);
    const double h = (XN - X1) / (N - 1);
    x.resize(N);
    std::vector<double> y(N);
    for (int i = 0; i < N; ++i) {
      x[i] = i * h + X1;
      y[i] = pot(x[i]);
    }
    spline(x, y, S1, Sn, a, b, c, d);
    const int np = (int)(pmax / hp + 1.5);
    P.resize(np);
    for (int i = 0; i < np; ++i) P[i] = i * hp;
  }
};

// The sums of phase two for momenta [i0, i1), on host threads.
inline void ftv_sums(const FtvSetup& s, std::vector<double>& Vpp) {
  const int N = FtvSetup::N;
  const std::vector<double>&x = s.x, &a = s.a, &b = s.b, &c = s.c, &d = s.d, &P = s.P;
  const double rc = s.rc, fc = s.fc, Qpi = 4.0 * kPi;
  const int np = (int)P.size();
  Vpp.assign(np, 0.0);
  auto work = [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      const double p = P[i];
      const double q = 2.0 * kPi * p;
      if (p == 0.0) {
        const double vp = std::pow(rc, 3) * fc / 3.0;
        double Fi = 0.0;
        for (int j = 0; j < N - 1; ++j) {
          const double u = x[j], w = x[j + 1];
          Fi = Fi + (a[j] / 3.0 * (std::pow(w, 3) - std::pow(u, 3)) + b[j] / 4.0 * (std::pow(w, 4) - std::pow(u, 4)) +
                     c[j] / 5.0 * (std::pow(w, 5) - std::pow(u, 5)) + d[j] / 6.0 * (std::pow(w, 6) - std::pow(u, 6)));
        }
        Vpp[i] = (vp + Fi) * Qpi;
        continue;
      }
      const double c1 = 1.0 / q, c2 = c1 * c1, c3 = c2 * c1, c4 = c2 * c2, c5 = c4 * c1;
      const double vp = fc * c1 * (std::sin(q * rc) * c1 - rc * std::cos(q * rc));
      double Fi = 0.0;
      double sj = std::sin(q * x[0]), cj = std::cos(q * x[0]);
      for (int j = 0; j < N - 1; ++j) {
        const double rj = x[j], rj1 = x[j + 1];
        const double sj1 = std::sin(q * rj1), cj1 = std::cos(q * rj1);
        const double A = a[j], B = b[j], C = c[j], D = d[j];
        Fi = Fi + ((A + (B + (C + D * rj) * rj) * rj) * rj * c1 * cj -
                   (A + (B + (C + D * rj1) * rj1) * rj1) * rj1 * c1 * cj1 +
                   (A + (2.0 * B + (3.0 * C + 4.0 * D * rj1) * rj1) * rj1) * c2 * sj1 -
                   (A + (2.0 * B + (3.0 * C + 4.0 * D * rj) * rj) * rj) * c2 * sj +
                   (2.0 * B + (6.0 * C + 12.0 * D * rj1) * rj1) * c3 * cj1 -
                   (2.0 * B + (6.0 * C + 12.0 * D * rj) * rj) * c3 * cj +
                   (6.0 * C + 24.0 * D * rj) * c4 * sj - (6.0 * C + 24.0 * D * rj1) * c4 * sj1 +
                   24.0 * D * c5 * (cj - cj1));
        sj = sj1;
        cj = cj1;
      }
      Vpp[i] = (vp + Fi) * c1 * Qpi;
    }
  };
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  const int chunk = (np + (int)nt - 1) / (int)nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int i0 = (int)t * chunk, i1 = std::min(np, i0 + chunk);
    if (i0 < i1) th.emplace_back(work, i0, i1);
  }
  for (auto& t : th) t.join();
}

struct FtvSpline {
  std::vector<double> P, A, B, C, D;

  FtvSpline(const MorsePot& pot, double pmax, double hp) {
    const FtvSetup s(pot, pmax, hp);
    std::vector<double> Vpp;
    ftv_sums(s, Vpp);
    finish(s.P, Vpp);
  }
  FtvSpline(const std::vector<double>& p, std::vector<double> Vpp) { finish(p, std::move(Vpp)); }

  void finish(const std::vector<double>& p, std::vector<double> Vpp) {
    const double small_q = 1e-2;
    P = p;
    const int np = (int)P.size();
    // small-q redefinition (FT_V_spline.f90:230-264)
    std::vector<double> xs{0.0}, ys{Vpp[0]};
    for (int i = 1; i < np && xs.size() < 4; ++i) {
      if (P[i] > small_q) {
        xs.push_back(P[i]);
        ys.push_back(Vpp[i]);
      }
    }
    std::vector<double> a4, b4, c4, d4;
    spline(xs, ys, 0.0, 0.0, a4, b4, c4, d4);
    for (int i = 1; i < np; ++i) {
      const double q = P[i];
      if (q <= small_q)
        Vpp[i] = a4[0] + q * (b4[0] + q * (c4[0] + q * d4[0]));
      else
        break;
    }
    spline(P, Vpp, 0.0, 0.0, A, B, C, D);
  }

  double operator()(double p) const {
    const int k = findi(P, p);
    return A[k] + p * (B[k] + p * (C[k] + p * D[k]));
  }
};

}  // namespace kinit
}  // namespace hetd
