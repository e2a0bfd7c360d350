// C++ drop-in layer tests (include/azeban_hip.hpp), written in the style of the
// reference's Catch2 tests: FFT vs brute-force DFT (test/azeban/operations/fft.cpp:
// 35-59,219-280) through make_fft with an owned and a shared work area, Taylor-Green
// steady state, the SSP-RK3 linear closed form (tests the same property as
// test/azeban/equations/spectral_viscosity.cpp:49-104, tracer-free) driven through
// TimeIntegrator<Dim>* / Equation<Dim>* base pointers, the reference's generic RK3
// sequence for a user-defined Equation against the fused step, the simulate_until
// clamp and both Simulation constructors.
#include <azeban_hip.hpp>

#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

static int g_fail = 0, g_checks = 0;
#define REQUIRE(cond)                                                    \
  do {                                                                   \
    ++g_checks;                                                          \
    if (!(cond)) {                                                       \
      ++g_fail;                                                          \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
    }                                                                    \
  } while (0)

using azeban::complex_t;
using azeban::real_t;
static const double pi = 3.14159265358979323846;

// separable brute-force forward DFT of a real (n^dim) array, last axis first
static std::vector<std::complex<double>> brute_r2c(const std::vector<double>& u, int dim, int n) {
  long total = 1;
  for (int d = 0; d < dim; ++d) total *= n;
  std::vector<std::complex<double>> a(u.begin(), u.end()), b(total);
  for (int ax = dim - 1; ax >= 0; --ax) {
    long stride = 1;
    for (int d = ax + 1; d < dim; ++d) stride *= n;
    for (long base = 0; base < total; ++base) {
      if ((base / stride) % n != 0) continue;
      for (int k = 0; k < n; ++k) {
        std::complex<double> s = 0;
        for (int x = 0; x < n; ++x)
          s += a[base + x * stride] * std::polar(1.0, -2 * pi * double(k) * x / n);
        b[base + k * stride] = s;
      }
    }
    a.swap(b);
  }
  return a;
}

template <int Dim>
static void test_fft(int n, int nvars) {
  azeban::shape_t<Dim + 1> rs, cs;
  rs[0] = cs[0] = nvars;
  for (int i = 1; i <= Dim; ++i) rs[i] = cs[i] = n;
  cs[Dim] = n / 2 + 1;
  azeban::device_array<real_t, Dim + 1> u(rs);
  azeban::device_array<complex_t, Dim + 1> uh(cs);
  std::mt19937 rng;
  std::uniform_real_distribution<double> dist(-1, 1);
  std::vector<double> h(u.size());
  for (auto& v : h) v = dist(rng);
  u.copy_to_device(h);
  // FFT<Dim>* from the factory, owning its work area (fft_factory.hpp:27-31)
  std::shared_ptr<azeban::FFT<Dim>> fft = azeban::make_fft<Dim>(uh.view(), u.view());
  REQUIRE(fft->is_forward() && fft->is_backward());
  REQUIRE(fft->shape_u_hat(u.shape()) == uh.shape() && fft->shape_u(uh.shape()) == u.shape());
  fft->forward();
  auto hh = uh.copy_to_host();
  long per = h.size() / nvars, perc = hh.size() / nvars;
  double err = 0;
  for (int v = 0; v < nvars; ++v) {
    std::vector<double> one(h.begin() + v * per, h.begin() + (v + 1) * per);
    auto ref = brute_r2c(one, Dim, n);
    for (long i = 0; i < perc; ++i) {
      long k = i % (n / 2 + 1), rest = i / (n / 2 + 1);
      auto r = ref[rest * n + k];
      auto g = hh[v * perc + i];
      err = std::max(err, std::max(std::abs(g.x - r.real()), std::abs(g.y - r.imag())));
    }
  }
  REQUIRE(err <= 1e-8);
  // a second transform planned without a work area, given a shared one
  // (FFT::initialize(..., false) + get_work_area_size / set_work_area, fft.hpp:86-90)
  auto fft2 = azeban::make_fft<Dim, real_t>(azeban::device_type::hip, azeban::FFT_BACKWARD);
  fft2->initialize(uh.view(), u.view(), false);
  REQUIRE(!fft2->is_forward() && fft2->is_backward());
  REQUIRE(fft2->get_work_area_size() == fft->get_work_area_size());
  fft2->set_work_area(fft->get_work_area());
  fft2->backward();
  auto back = u.copy_to_host();
  double nd = std::pow(double(n), Dim), err2 = 0;
  for (size_t i = 0; i < h.size(); ++i) err2 = std::max(err2, std::abs(back[i] / nd - h[i]));
  REQUIRE(err2 <= 1e-8);
  std::printf("FFT %dD n=%d n_vars=%d: max err %.2e, round trip %.2e\n", Dim, n, nvars, err, err2);
}

// physical field -> Fourier state: FFT then Leray (Initializer::initialize)
static azeban::device_array<complex_t, 3> init2d(int N, const std::vector<double>& phys) {
  azeban::device_array<real_t, 3> u({2, N, N});
  u.copy_to_device(phys);
  azeban::device_array<complex_t, 3> uh({2, N, N / 2 + 1});
  azeban::make_fft<2>(uh.view(), u.view(), azeban::FFT_FORWARD)->forward();
  azeban::leray<2>(uh);
  return uh;
}

static void test_taylor_green_steady() {
  const int N = 128;
  std::vector<double> p(2 * N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double x = double(i) / N, y = double(j) / N;
      p[i * N + j] = std::sin(2 * pi * x) * std::cos(2 * pi * y);
      p[N * N + i * N + j] = -std::cos(2 * pi * x) * std::sin(2 * pi * y);
    }
  auto uh = init2d(N, p);
  azeban::Grid<2> grid(N);
  auto eq = std::make_shared<azeban::IncompressibleEuler<2>>(grid, azeban::SmoothCutoff1D(0.05 / N, 1, 2 * pi * N / 4));
  azeban::device_array<complex_t, 3> d(uh.shape());
  eq->dudt(d, uh, 0, 0, 0.2);
  auto hd = d.copy_to_host();
  auto hu = uh.copy_to_host();
  double md = 0, mu = 0;
  for (auto& z : hd) md = std::max(md, std::hypot(z.x, z.y));
  for (auto& z : hu) mu = std::max(mu, std::hypot(z.x, z.y));
  REQUIRE(md <= 1e-11 * mu);
  REQUIRE(std::abs(eq->dt(0.2) - 0.2 / (N * std::pow(2.0 / 3.0, 2))) <= 1e-13);
  std::printf("Taylor-Green 2D N=%d: max|dudt|/max|u| = %.2e\n", N, md / mu);
}

static void test_rk3_linear_closed_form() {
  // shear flow u = (sin(2 pi m y), 0): (u.grad)u = 0, each step multiplies by R(nu dt)
  const int N = 64, m = 20;
  std::vector<double> p(2 * N * N, 0.0);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) p[i * N + j] = std::sin(2 * pi * m * double(j) / N);
  auto uh0 = init2d(N, p);
  azeban::Grid<2> grid(N);
  azeban::SmoothCutoff1D visc(0.05 / N, 1, 2 * pi * N / 4);
  // the reference's construction chain, held through the base classes
  std::shared_ptr<azeban::Equation<2>> eq = std::make_shared<azeban::IncompressibleEuler<2>>(grid, visc);
  std::shared_ptr<azeban::TimeIntegrator<2>> rk
      = std::make_shared<azeban::SSP_RK3<2>>(azeban::device_type::hip, grid.shape_fourier(2), eq);
  azeban::Simulation<2> sim(uh0.view(), grid, 0.2, rk);
  REQUIRE(sim.equation() == eq && rk->equation() == eq && sim.memory_location() == azeban::device_type::hip);
  auto h0 = uh0.copy_to_host();
  const double nu = visc.eval(2 * pi * m);
  double ratio = 1.0, t = 0;
  for (int s = 0; s < 5; ++s) {
    const double max_dt = std::min(1.0 - t, 0.2 * 2. / (eq->visc() * (pi * N) * (pi * N)));
    const double dt = rk->integrate(1.0, max_dt, 0.2, sim.u());
    const double z = nu * dt;
    ratio *= 1 + z + z * z / 2 + z * z * z / 6;
    t += dt;
  }
  auto h = azeban::device_array<complex_t, 3>::to_host(sim.u());
  const long idx = m;  // component 0, row 0, column m
  const double expect_x = h0[idx].x * ratio, expect_y = h0[idx].y * ratio;
  // sin gives a purely imaginary mode: compare against the mode's magnitude
  const double mag = std::hypot(h0[idx].x, h0[idx].y);
  REQUIRE(std::abs(h[idx].x - expect_x) <= 1e-12 * mag);
  REQUIRE(std::abs(h[idx].y - expect_y) <= 1e-12 * mag);
  std::printf("RK3 linear closed form: ratio %.15f\n", ratio);
}

// A user-defined Equation (not IncompressibleEuler) takes SSP_RK3's generic path: the
// reference's three dudt calls and five axpbys (ssp_rk3.hpp:48-64) on device arrays.
// Wrapping the MI355X equation, it must reproduce the fused step to round-off.
template <int Dim>
class Wrapped final : public azeban::Equation<Dim> {
 public:
  explicit Wrapped(std::shared_ptr<azeban::Equation<Dim>> inner) : azeban::Equation<Dim>(Grid0()), in_(inner) {}
  void dudt(const azeban::array_view<complex_t, Dim + 1>& d, const azeban::array_const_view<complex_t, Dim + 1>& u,
            double t, double dt, double C) override {
    ++calls;
    in_->dudt(d, u, t, dt, C);
  }
  double dt(double C) const override { return in_->dt(C); }
  int n_vars() const override { return in_->n_vars(); }
  real_t visc() const override { return in_->visc(); }
  int calls = 0;

 private:
  static azeban::Grid<Dim> Grid0() { return azeban::Grid<Dim>(); }
  std::shared_ptr<azeban::Equation<Dim>> in_;
};

static void test_generic_rk3_matches_fused() {
  const int N = 64;
  std::vector<double> p(2 * N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double y = double(j) / N, x = double(i) / N;
      p[i * N + j] = y <= 0.5 ? std::tanh((y - 0.25) / 0.05) : std::tanh((0.75 - y) / 0.05);
      p[N * N + i * N + j] = 0.05 * std::sin(2 * pi * x);
    }
  auto uh = init2d(N, p);
  azeban::Grid<2> grid(N);
  azeban::SmoothCutoff1D visc(0.05 / N, 1, 2 * pi * N / 4);
  auto fused_eq = std::make_shared<azeban::IncompressibleEuler<2>>(grid, visc);
  std::shared_ptr<azeban::TimeIntegrator<2>> fused
      = std::make_shared<azeban::SSP_RK3<2>>(azeban::device_type::hip, grid.shape_fourier(2), fused_eq);
  auto wrapped = std::make_shared<Wrapped<2>>(std::make_shared<azeban::IncompressibleEuler<2>>(grid, visc));
  std::shared_ptr<azeban::TimeIntegrator<2>> generic
      = std::make_shared<azeban::SSP_RK3<2>>(azeban::device_type::hip, grid.shape_fourier(2), wrapped);
  azeban::Simulation<2> a(uh.view(), grid, 0.2, fused), b(uh.view(), grid, 0.2, generic);
  a.simulate_until(0.05);
  b.simulate_until(0.05);
  REQUIRE(a.time() == 0.05 && b.time() == 0.05);
  REQUIRE(a.steps() == b.steps() && wrapped->calls == 3 * b.steps());
  auto ha = azeban::device_array<complex_t, 3>::to_host(a.u());
  auto hb = azeban::device_array<complex_t, 3>::to_host(b.u());
  double num = 0, den = 0;
  for (size_t i = 0; i < ha.size(); ++i) {
    num += std::pow(ha[i].x - hb[i].x, 2) + std::pow(ha[i].y - hb[i].y, 2);
    den += ha[i].x * ha[i].x + ha[i].y * ha[i].y;
  }
  REQUIRE(std::sqrt(num / den) <= 1e-13);
  std::printf("generic SSP_RK3 (dudt + axpby) vs fused step: %ld steps, rel L2 %.1e\n", b.steps(),
              std::sqrt(num / den));
}

static void test_simulate_until_clamp() {
  const int N = 64;
  std::vector<double> p(2 * N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double y = double(j) / N, x = double(i) / N;
      p[i * N + j] = y <= 0.5 ? std::tanh((y - 0.25) / 0.05) : std::tanh((0.75 - y) / 0.05);
      p[N * N + i * N + j] = 0.05 * std::sin(2 * pi * x);
    }
  auto uh = init2d(N, p);
  azeban::Grid<2> grid(N);
  azeban::SmoothCutoff1D visc(0.05 / N, 1, 2 * pi * N / 4);
  auto eq = std::make_shared<azeban::IncompressibleEuler<2>>(grid, visc);
  auto rk = std::make_shared<azeban::SSP_RK3<2>>(azeban::device_type::hip, grid.shape_fourier(2), eq);
  azeban::Simulation<2> a(grid, 0.2, rk, azeban::device_type::hip);
  azeban::copy<complex_t, 3>(a.u(), uh.view());
  a.simulate_until(0.05);
  auto eq2 = std::make_shared<azeban::IncompressibleEuler<2>>(grid, visc);
  auto rk2 = std::make_shared<azeban::SSP_RK3<2>>(azeban::device_type::hip, grid.shape_fourier(2), eq2);
  azeban::Simulation<2> b(grid, 0.2, rk2, azeban::device_type::hip);
  azeban::copy<complex_t, 3>(b.u(), uh.view());
  b.simulate_until_device(0.05, 4);
  REQUIRE(a.time() == 0.05);
  REQUIRE(b.time() == 0.05);
  REQUIRE(a.steps() == b.steps());
  auto ha = azeban::device_array<complex_t, 3>::to_host(a.u());
  auto hb = azeban::device_array<complex_t, 3>::to_host(b.u());
  double num = 0, den = 0;
  for (size_t i = 0; i < ha.size(); ++i) {
    num += std::pow(ha[i].x - hb[i].x, 2) + std::pow(ha[i].y - hb[i].y, 2);
    den += ha[i].x * ha[i].x + ha[i].y * ha[i].y;
  }
  REQUIRE(std::sqrt(num / den) <= 1e-14);
  auto e = azeban::energy_spectrum<2>(grid, a.u());
  REQUIRE((long)e.size() == grid.N_fourier);
  std::printf("simulate_until: %ld host-loop steps, device loop agrees to %.1e\n", a.steps(), std::sqrt(num / den));
}

int main() {
  test_fft<1>(128, 1);
  test_fft<1>(128, 2);
  test_fft<2>(128, 1);
  test_fft<2>(128, 2);
  test_fft<3>(32, 1);
  test_fft<3>(32, 2);
  test_taylor_green_steady();
  test_rk3_linear_closed_form();
  test_generic_rk3_matches_fused();
  test_simulate_until_clamp();
  std::printf("%d checks, %d failed\n", g_checks, g_fail);
  return g_fail == 0 ? 0 : 1;
}
