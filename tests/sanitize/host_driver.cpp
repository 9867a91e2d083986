// Host-side sanitizer driver: calls the C-ABI through its host code paths
// (argument checks, the k-space generators, context set-up, steps, the
// batch executor's threads, checkpoints, tuning, device-built starts) in a
// build whose host code carries AddressSanitizer and
// UndefinedBehaviorSanitizer (device code is built as in the library).
//   host_driver cpu   host-only paths (no GPU needed)
//   host_driver gpu   everything, on GPU 0
// Exit status: the number of failed checks (sanitizer reports abort).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/hetddft.h"

static int g_fail = 0;
#define EXPECT(cond, ...)                        \
  do {                                           \
    if (!(cond)) {                               \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                  \
      std::printf("\n");                         \
      ++g_fail;                                  \
    }                                            \
  } while (0)
#define OK(call, ctx) EXPECT((call) == HETDDFT_OK, "%s: %s", #call, hetddft_last_error(ctx))

static hetddft_params params(int n, double xmax, int npar, int precision) {
  hetddft_params p;
  hetddft_default_params(&p);
  p.nx = p.ny = p.nz = n;
  p.xmax = p.ymax = p.zmax = xmax;
  p.n_par = npar;
  p.precision = precision;
  return p;
}

static hetddft_morse morse() {
  hetddft_morse m;
  m.D = 20.0;
  m.re = 5.0;
  m.alpha = 0.8;
  m.rc = 2.0;
  m.umax = 0.0;
  m.nq = 1500;
  return m;
}

static size_t half(const hetddft_params& p) { return (size_t)(p.nx / 2 + 1) * p.ny * p.nz; }

static void cpu_paths() {
  hetddft_params p;
  EXPECT(hetddft_default_params(&p) == HETDDFT_OK, "default_params");
  EXPECT(hetddft_default_params(nullptr) == HETDDFT_E_ARG, "default_params(NULL)");
  hetddft_ctx* c = nullptr;
  hetddft_params bad = params(7, 8.0, 16, HETDDFT_FP64);  // odd nx
  EXPECT(hetddft_create(&bad, &c) == HETDDFT_E_ARG && c == nullptr, "create(odd nx)");
  bad = params(26, 8.0, 16, HETDDFT_FP64);  // not in the tstgrid list
  EXPECT(hetddft_create(&bad, &c) == HETDDFT_E_ARG, "create(26)");
  EXPECT(hetddft_create(nullptr, &c) == HETDDFT_E_ARG, "create(NULL)");
  EXPECT(hetddft_destroy(nullptr) == HETDDFT_OK, "destroy(NULL) is a no-op");
  EXPECT(hetddft_last_error(nullptr) != nullptr, "last_error(NULL)");
  hetddft_batch* b = nullptr;
  EXPECT(hetddft_batch_create(nullptr, 1, 1, &b) == HETDDFT_E_ARG, "batch_create(NULL)");
  EXPECT(hetddft_batch_destroy(nullptr) == HETDDFT_OK, "batch_destroy(NULL) is a no-op");
  const hetddft_morse m = morse();
  for (int n : {24, 32, 48}) {
    for (int prec : {HETDDFT_FP64, HETDDFT_MIXED}) {
      const hetddft_params q = params(n, 0.5 * n * 0.5, 64, prec);
      std::vector<double> f(half(q)), w(half(q)), k(half(q)), v(half(q));
      EXPECT(hetddft_make_kernels(&q, &m, f.data(), w.data(), k.data(), v.data()) == HETDDFT_OK, "make_kernels %d", n);
      bool finite = true;
      for (size_t i = 0; i < f.size(); ++i)
        finite = finite && std::isfinite(f[i]) && std::isfinite(w[i]) && std::isfinite(k[i]) && std::isfinite(v[i]);
      EXPECT(finite, "kernels finite at %d", n);
    }
  }
  EXPECT(hetddft_make_kernels(nullptr, &m, nullptr, nullptr, nullptr, nullptr) == HETDDFT_E_ARG, "make_kernels(NULL)");
}

static void gpu_paths() {
  const int grids[5] = {24, 24, 24, 32, 32};
  std::vector<hetddft_ctx*> ctx(5, nullptr);
  std::vector<double> dts(5, 6.546e-5);
  const hetddft_morse m = morse();
  for (int i = 0; i < 5; ++i) {
    const int n = grids[i];
    const hetddft_params p = params(n, 0.25 * n, 128, i % 2 ? HETDDFT_MIXED : HETDDFT_FP64);
    EXPECT(hetddft_create(&p, &ctx[i]) == HETDDFT_OK, "create %d", i);
    if (!ctx[i]) return;
    hetddft_ctx* c = ctx[i];
    if (i < 3) {
      std::vector<double> f(half(p)), w(half(p)), k(half(p)), v(half(p));
      OK(hetddft_make_kernels(&p, &m, f.data(), w.data(), k.data(), v.data()), c);
      OK(hetddft_set_kernels(c, f.data(), w.data(), k.data(), v.data()), c);
    } else {
      OK(hetddft_init_kernels(c, &m), c);
    }
    hetddft_droplet d;
    std::memset(&d, 0, sizeof d);
    d.n_he = 12.0;
    d.rho0 = 0.0218;
    d.r0 = 2.22;
    d.a = 0.7;
    d.cavity = 2.0;
    d.cavity_w = 0.3;
    d.ripple = i == 4 ? 0.2 : 0.0;
    for (int k = 0; k < 24; ++k) d.clm[k] = 0.1 * (k % 5) - 0.2;
    d.phase_noise = 0.01;
    d.sigma = 0.5;
    d.v0[2] = 1.0;
    d.seed = 100 + i;
    OK(hetddft_init_droplet(c, &d), c);
    OK(hetddft_start(c), c);
    double err[3];
    for (int s = 0; s < 5; ++s) OK(hetddft_iterate(c, dts[i], err), c);
    double means[9], en[12], diag[HETDDFT_DIAG_N];
    OK(hetddft_particle_means(c, means), c);
    OK(hetddft_energy(c, en), c);
    OK(hetddft_diagnostics(c, diag), c);
    std::vector<double> psi(2 * (size_t)n * n * n), r(3 * 128), v(3 * 128), a(3 * 128), den((size_t)n * n * n);
    OK(hetddft_get_state(c, psi.data(), r.data(), v.data(), a.data()), c);
    OK(hetddft_get_field(c, 0, den.data()), c);
    EXPECT(hetddft_get_field(c, 999, den.data()) == HETDDFT_E_ARG, "get_field(999)");
  }
  // checkpoint round trip, then continue
  long long nb = 0;
  OK(hetddft_checkpoint_size(ctx[0], &nb), ctx[0]);
  std::vector<unsigned char> img((size_t)nb);
  OK(hetddft_checkpoint_save(ctx[0], img.data()), ctx[0]);
  double e1[3], e2[3];
  OK(hetddft_iterate(ctx[0], dts[0], e1), ctx[0]);
  OK(hetddft_checkpoint_load(ctx[0], img.data()), ctx[0]);
  OK(hetddft_iterate(ctx[0], dts[0], e2), ctx[0]);
  EXPECT(e1[0] == e2[0] && e1[1] == e2[1] && e1[2] == e2[2], "resume from checkpoint");
  // tuning
  hetddft_tuning t;
  OK(hetddft_autotune(ctx[3], dts[3], 1, &t), ctx[3]);
  OK(hetddft_get_tuning(ctx[3], &t), ctx[3]);
  OK(hetddft_set_tuning(ctx[3], &t), ctx[3]);
  t.tile[0] = 9;
  EXPECT(hetddft_set_tuning(ctx[3], &t) == HETDDFT_E_ARG, "set_tuning(tile 9)");
  // sorteo from a Gaussian density
  {
    const int n = 24;
    std::vector<double> dx((size_t)n * n * n);
    const double h = 2.0 * 0.25 * n / n;
    for (int z = 0; z < n; ++z)
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
          const double X = -6.0 + h * x, Y = -6.0 + h * y, Z = -6.0 + h * z;
          dx[x + (size_t)n * (y + (size_t)n * z)] = std::exp(-(X * X + Y * Y + Z * Z));
        }
    const double lo[3] = {-3, -3, -3}, hi[3] = {3, 3, 3};
    std::vector<double> rimp(3 * 128);
    OK(hetddft_sample_particles(ctx[1], dx.data(), lo, hi, 7, 1 << 20, rimp.data()), ctx[1]);
  }
  // the batch executor on 3 host threads (two grid groups, both precisions)
  hetddft_batch* b = nullptr;
  OK(hetddft_batch_create(ctx.data(), 5, 3, &b), nullptr);
  if (b) {
    std::vector<double> errs(15);
    int failed = -2;
    EXPECT(hetddft_batch_iterate(b, dts.data(), 4, errs.data(), &failed) == HETDDFT_OK && failed == -1,
           "batch_iterate: %s", hetddft_batch_last_error(b));
    EXPECT(hetddft_batch_iterate(b, dts.data(), 3, nullptr, nullptr) == HETDDFT_OK, "batch_iterate(no errors)");
    OK(hetddft_batch_destroy(b), nullptr);
  }
  hetddft_ctx* dup[2] = {ctx[0], ctx[0]};
  EXPECT(hetddft_batch_create(dup, 2, 1, &b) == HETDDFT_E_ARG, "batch_create(duplicate)");
  for (hetddft_ctx* c : ctx) OK(hetddft_destroy(c), nullptr);
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  std::printf("host sanitizer driver: %s paths\n", gpu ? "cpu and gpu" : "cpu");
  std::fflush(stdout);
  cpu_paths();
  if (gpu) gpu_paths();
  std::printf("%d failed checks\n", g_fail);
  return g_fail;
}
