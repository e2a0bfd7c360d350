// azeban_hip.hpp -- C++ host layer over the C ABI (azeban_hip.h), mirroring
// TobiasRohner/azeban's Equation / FFT / TimeIntegrator / Simulation interface for
// the IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> + SSP_RK3 path.
//
// Mirrored reference interfaces (paths relative to the reference repository):
//   Grid<Dim>                    include/azeban/grid.hpp:30-165
//   SmoothCutoff1D               include/azeban/equations/spectral_viscosity.hpp:56-77
//   Equation<Dim>                include/azeban/equations/equation.hpp:27-56
//   IncompressibleEuler          include/azeban/equations/incompressible_euler.hpp:39-124
//   FFT<Dim> (R2C)               include/azeban/operations/fft.hpp:37-121
//   TimeIntegrator<Dim>          include/azeban/evolution/time_integrator.hpp:30-60
//   SSP_RK3<Dim>                 include/azeban/evolution/ssp_rk3.hpp:28-64
//   Simulation<Dim>              include/azeban/simulation.hpp:30-93, src/azeban/simulation.cpp:89-102
//
// Error behaviour follows the reference: a failing call prints the message and
// exits (include/azeban/logging.hpp:24-32).  Arrays are device arrays (the `hip`
// memory location); host arrays are copied with copy_to_host / copy_to_device.
#ifndef AZEBAN_HIP_HPP_
#define AZEBAN_HIP_HPP_

#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "azeban_hip.h"

namespace azeban {

using real_t = double;
struct complex_t {
  real_t x, y;
};

inline void check(int rc, const char* what) {
  if (rc != AZB_OK) {
    std::fprintf(stderr, "Error in %s: %s\n", what, azb_last_error());
    std::exit(1);
  }
}

enum class device_type { cpu, hip };

template <int Rank>
using shape_t = std::array<long, Rank>;

template <int Rank>
long product(const shape_t<Rank>& s) {
  long p = 1;
  for (long v : s) p *= v;
  return p;
}

// Non-owning view of a row-major device array (the zisa::array_view role).
template <typename T, int Rank>
class array_view {
 public:
  array_view() = default;
  array_view(const shape_t<Rank>& shape, T* ptr) : shape_(shape), ptr_(ptr) {}
  T* raw() const { return ptr_; }
  long shape(int i) const { return shape_[i]; }
  const shape_t<Rank>& shape() const { return shape_; }
  long size() const { return product<Rank>(shape_); }
  device_type memory_location() const { return device_type::hip; }

 private:
  shape_t<Rank> shape_{};
  T* ptr_ = nullptr;
};

template <typename T, int Rank>
using array_const_view = array_view<const T, Rank>;

// Owning device array (the zisa::array role, device memory only).
template <typename T, int Rank>
class device_array {
 public:
  explicit device_array(const shape_t<Rank>& shape) : shape_(shape) {
    void* p = nullptr;
    check(azb_malloc(&p, sizeof(T) * product<Rank>(shape)), "azb_malloc");
    ptr_.reset(static_cast<T*>(p));
  }
  T* raw() const { return ptr_.get(); }
  long size() const { return product<Rank>(shape_); }
  const shape_t<Rank>& shape() const { return shape_; }
  long shape(int i) const { return shape_[i]; }
  array_view<T, Rank> view() const { return {shape_, ptr_.get()}; }
  operator array_view<T, Rank>() const { return view(); }
  operator array_view<const T, Rank>() const { return {shape_, ptr_.get()}; }
  void copy_to_device(const std::vector<T>& h) {
    check(azb_memcpy_h2d(raw(), h.data(), sizeof(T) * size()), "azb_memcpy_h2d");
  }
  std::vector<T> copy_to_host() const {
    std::vector<T> h(size());
    check(azb_memcpy_d2h(h.data(), raw(), sizeof(T) * size()), "azb_memcpy_d2h");
    return h;
  }

 private:
  struct Del {
    void operator()(T* p) const { azb_free(p); }
  };
  shape_t<Rank> shape_;
  std::unique_ptr<T, Del> ptr_;
};

template <int Dim>
struct Grid {
  static constexpr int dim_v = Dim;
  long N_phys = 0, N_phys_pad = 0, N_fourier = 0, N_fourier_pad = 0;
  Grid() = default;
  explicit Grid(long N) : Grid(N, 3 * N / 2) {}
  Grid(long N, long M) : N_phys(N), N_phys_pad(M), N_fourier(N / 2 + 1), N_fourier_pad(M / 2 + 1) {}
  shape_t<Dim + 1> shape_fourier(long n_vars) const {
    shape_t<Dim + 1> s;
    s[0] = n_vars;
    for (int i = 1; i < Dim; ++i) s[i] = N_phys;
    s[Dim] = N_fourier;
    return s;
  }
  shape_t<Dim + 1> shape_phys(long n_vars) const {
    shape_t<Dim + 1> s;
    s[0] = n_vars;
    for (int i = 1; i <= Dim; ++i) s[i] = N_phys;
    return s;
  }
};

struct SmoothCutoff1D {
  real_t eps, s, k0;
  SmoothCutoff1D(real_t eps_, real_t s_, real_t k0_) : eps(eps_), s(s_), k0(k0_) {}
  real_t Qk(real_t k) const {
    const real_t k1 = std::abs(k) / k0, k2 = k1 * k1, k4 = k2 * k2, k8 = k4 * k4, k16 = k8 * k8;
    return 1 - std::exp(-(k16 * k2));
  }
  real_t eval(real_t k) const { return -eps * std::pow(k, 2 * s) * Qk(k); }
};

template <int Dim>
class Equation {
 public:
  explicit Equation(const Grid<Dim>& grid) : grid_(grid) {}
  virtual ~Equation() = default;
  virtual void dudt(const array_view<complex_t, Dim + 1>& dudt_hat,
                    const array_const_view<complex_t, Dim + 1>& u_hat, double t, double dt, double C) = 0;
  virtual double dt(double C) const = 0;
  virtual int n_vars() const = 0;
  virtual real_t visc() const { return 0; }
  const Grid<Dim>& grid() const { return grid_; }

 protected:
  Grid<Dim> grid_;
};

enum class precision { fp64 = AZB_FP64, fp32_mixed = AZB_FP32_MIXED };

// IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> on the MI355X.
template <int Dim>
class IncompressibleEuler final : public Equation<Dim> {
 public:
  IncompressibleEuler(const Grid<Dim>& grid, const SmoothCutoff1D& visc, device_type device = device_type::hip,
                      precision prec = precision::fp64, void* stream = nullptr)
      : Equation<Dim>(grid), visc_(visc) {
    if (device != device_type::hip) {
      std::fprintf(stderr, "Error: IncompressibleEuler on the MI355X needs device_type::hip\n");
      std::exit(1);
    }
    azb_solver* h = nullptr;
    check(azb_create(Dim, (int)grid.N_phys, (int)grid.N_phys_pad, Dim, static_cast<int>(prec),
                     azb_sv_params{visc.eps, visc.s, visc.k0}, stream, &h),
          "azb_create");
    handle_.reset(h);
  }
  void dudt(const array_view<complex_t, Dim + 1>& dudt_hat, const array_const_view<complex_t, Dim + 1>& u_hat,
            double, double, double) override {
    if (dudt_hat.shape(0) != n_vars() || u_hat.shape(0) != n_vars()) {
      std::fprintf(stderr, "Error: Wrong number of variables\n");
      std::exit(1);
    }
    check(azb_dudt(handle_.get(), u_hat.raw(), dudt_hat.raw()), "azb_dudt");
    check(azb_synchronize(handle_.get()), "azb_synchronize");
  }
  double dt(double C) const override {
    double v = 0;
    check(azb_dt(handle_.get(), C, &v), "azb_dt");
    return v;
  }
  int n_vars() const override { return Dim; }
  real_t visc() const override { return visc_.eps; }
  azb_solver* handle() const { return handle_.get(); }

 private:
  struct Del {
    void operator()(azb_solver* s) const { azb_destroy(s); }
  };
  SmoothCutoff1D visc_;
  std::unique_ptr<azb_solver, Del> handle_;
};

template <int Dim>
class TimeIntegrator {
 public:
  TimeIntegrator(device_type device, const std::shared_ptr<IncompressibleEuler<Dim>>& eq)
      : device_(device), equation_(eq) {}
  virtual ~TimeIntegrator() = default;
  virtual double integrate(double t, double max_dt, double C, const array_view<complex_t, Dim + 1>& u) = 0;
  std::shared_ptr<IncompressibleEuler<Dim>> equation() const { return equation_; }
  device_type memory_location() const { return device_; }

 protected:
  device_type device_;
  std::shared_ptr<IncompressibleEuler<Dim>> equation_;
};

template <int Dim>
class SSP_RK3 final : public TimeIntegrator<Dim> {
 public:
  SSP_RK3(device_type device, const shape_t<Dim + 1>&, const std::shared_ptr<IncompressibleEuler<Dim>>& eq)
      : TimeIntegrator<Dim>(device, eq) {}
  double integrate(double, double max_dt, double C, const array_view<complex_t, Dim + 1>& u) override {
    double dt = 0;
    check(azb_rk3_step(this->equation_->handle(), u.raw(), max_dt, C, &dt), "azb_rk3_step");
    return dt;
  }
};

template <int Dim>
class Simulation {
 public:
  Simulation(const Grid<Dim>& grid, double C, const std::shared_ptr<SSP_RK3<Dim>>& ts, device_type)
      : u_(grid.shape_fourier(Dim)), C_(C), grid_(grid), timestepper_(ts) {}
  // host-driven loop, exactly Simulation::simulate_until (src/azeban/simulation.cpp:89-102)
  void simulate_until(double t) {
    const double pi = 3.14159265358979323846;
    while (time_ < t) {
      const real_t eps = timestepper_->equation()->visc();
      const double max_dt = std::min(t - time_, C_ * 2. / (eps * (pi * grid_.N_phys) * (pi * grid_.N_phys)));
      const double dt = timestepper_->integrate(t, max_dt, C_, u_.view());
      if (dt <= 1e-10) std::fprintf(stderr, "Warning: Timestep is tiny. dt = %g\n", dt);
      time_ += dt;
      ++steps_;
    }
    time_ = t;
  }
  // device-driven loop: dt and time stay on the device, the host checks every chunk steps
  void simulate_until_device(double t, int chunk = 8) {
    int steps = 0;
    check(azb_simulate_until(timestepper_->equation()->handle(), u_.raw(), t, C_, &time_, 0, chunk, &steps),
          "azb_simulate_until");
    steps_ += steps;
  }
  void simulate_for(double t) { simulate_until(time_ + t); }
  device_array<complex_t, Dim + 1>& u() { return u_; }
  double time() const { return time_; }
  void set_time(double t) { time_ = t; }
  void reset() { time_ = 0; }
  long steps() const { return steps_; }

 private:
  device_array<complex_t, Dim + 1> u_;
  double C_;
  Grid<Dim> grid_;
  std::shared_ptr<SSP_RK3<Dim>> timestepper_;
  double time_ = 0;
  long steps_ = 0;
};

// FFT<Dim> R2C on device arrays: forward() = unnormalised r2c, backward() = c2r
// (which, like FFTW's and cuFFT's, may overwrite u_hat).
template <int Dim>
class FFT {
 public:
  FFT(const array_view<complex_t, Dim + 1>& u_hat, const array_view<real_t, Dim + 1>& u) : u_hat_(u_hat), u_(u) {}
  void forward() {
    check(azb_fft_r2c(Dim, (int)u_.shape(1), (int)u_.shape(0), AZB_FP64, u_.raw(), u_hat_.raw(), nullptr),
          "azb_fft_r2c");
  }
  void backward() {
    check(azb_fft_c2r(Dim, (int)u_.shape(1), (int)u_.shape(0), AZB_FP64, u_hat_.raw(), u_.raw(), nullptr),
          "azb_fft_c2r");
  }
  const array_view<complex_t, Dim + 1>& u_hat() const { return u_hat_; }
  const array_view<real_t, Dim + 1>& u() const { return u_; }

 private:
  array_view<complex_t, Dim + 1> u_hat_;
  array_view<real_t, Dim + 1> u_;
};

template <int Dim>
void leray(const array_view<complex_t, Dim + 1>& u_hat) {
  check(azb_leray(Dim, (int)u_hat.shape(1), u_hat.raw(), nullptr), "azb_leray");
}

template <int Dim>
std::vector<real_t> energy_spectrum(const Grid<Dim>& grid, const array_view<complex_t, Dim + 1>& u_hat) {
  std::vector<real_t> e(grid.N_fourier);
  check(azb_energy_spectrum(Dim, (int)grid.N_phys, u_hat.raw(), e.data(), nullptr), "azb_energy_spectrum");
  return e;
}

}  // namespace azeban

#endif  // AZEBAN_HIP_HPP_
