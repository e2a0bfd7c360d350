// azeban_hip.hpp -- C++ host layer over the C ABI (azeban_hip.h), with the class
// shapes of TobiasRohner/azeban's Equation / FFT / TimeIntegrator / Simulation
// interface for the IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> + SSP_RK3
// path, so code written against those classes compiles against this header.
//
// Mirrored reference interfaces (paths relative to the reference repository):
//   Grid<Dim>                    include/azeban/grid.hpp:30-165
//   SmoothCutoff1D               include/azeban/equations/spectral_viscosity.hpp:56-77
//   Equation<Dim>                include/azeban/equations/equation.hpp:27-56
//   IncompressibleEuler          include/azeban/equations/incompressible_euler.hpp:39-124
//   FFT<Dim, ScalarU>            include/azeban/operations/fft.hpp:37-121
//   make_fft                     include/azeban/operations/fft_factory.hpp:27-60,
//                                src/azeban/operations/fft_factory.cpp:10-40
//   TimeIntegrator<Dim>          include/azeban/evolution/time_integrator.hpp:30-60
//   SSP_RK3<Dim>                 include/azeban/evolution/ssp_rk3.hpp:28-64
//   Simulation<Dim>              include/azeban/simulation.hpp:30-93, src/azeban/simulation.cpp:6-102
//
// The memory location `device_type::hip` stands where the reference has
// zisa::device_type::cuda: arrays there live in device memory obtained through
// azb_malloc (device_array), and views carry their location.
//
// Error behaviour follows the reference: a failing call prints the message and
// exits (include/azeban/logging.hpp:24-32).
#ifndef AZEBAN_HIP_HPP_
#define AZEBAN_HIP_HPP_

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <type_traits>
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

[[noreturn]] inline void fatal(const char* msg) {
  std::fprintf(stderr, "Error: %s\n", msg);
  std::exit(1);
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

// Non-owning view of a row-major array (the zisa::array_view role) with its memory
// location; views made from a device_array are `hip`.
template <typename T, int Rank>
class array_view {
 public:
  array_view() = default;
  array_view(const shape_t<Rank>& shape, T* ptr, device_type loc = device_type::hip)
      : shape_(shape), ptr_(ptr), loc_(loc) {}
  // a view of mutable data is also a view of const data
  template <typename U, typename = std::enable_if_t<std::is_same_v<const U, T> && !std::is_same_v<U, T>>>
  array_view(const array_view<U, Rank>& v) : shape_(v.shape()), ptr_(v.raw()), loc_(v.memory_location()) {}
  T* raw() const { return ptr_; }
  long shape(int i) const { return shape_[i]; }
  const shape_t<Rank>& shape() const { return shape_; }
  long size() const { return product<Rank>(shape_); }
  device_type memory_location() const { return loc_; }

 private:
  shape_t<Rank> shape_{};
  T* ptr_ = nullptr;
  device_type loc_ = device_type::hip;
};

template <typename T, int Rank>
using array_const_view = array_view<const T, Rank>;

// Owning device array (the zisa::array role for device memory).
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
  device_type memory_location() const { return device_type::hip; }
  array_view<T, Rank> view() const { return {shape_, ptr_.get()}; }
  operator array_view<T, Rank>() const { return view(); }
  operator array_view<const T, Rank>() const { return {shape_, ptr_.get()}; }
  void copy_to_device(const std::vector<T>& h) {
    if ((long)h.size() != size()) fatal("copy_to_device: size mismatch");
    check(azb_memcpy_h2d(raw(), h.data(), sizeof(T) * size()), "azb_memcpy_h2d");
  }
  std::vector<T> copy_to_host() const { return to_host(view()); }

  static std::vector<T> to_host(const array_view<const T, Rank>& v) {
    std::vector<T> h(v.size());
    check(azb_memcpy_d2h(h.data(), v.raw(), sizeof(T) * v.size()), "azb_memcpy_d2h");
    return h;
  }

 private:
  struct Del {
    void operator()(T* p) const { azb_free(p); }
  };
  shape_t<Rank> shape_;
  std::unique_ptr<T, Del> ptr_;
};

// copy between views of equal size (zisa::copy for the locations used here)
template <typename T, int Rank>
void copy(const array_view<T, Rank>& dst, const array_view<const T, Rank>& src) {
  if (dst.size() != src.size()) fatal("copy: size mismatch");
  const size_t bytes = sizeof(T) * dst.size();
  const bool dh = dst.memory_location() == device_type::hip, sh = src.memory_location() == device_type::hip;
  if (dh && sh)
    check(azb_memcpy_d2d(dst.raw(), src.raw(), bytes), "azb_memcpy_d2d");
  else if (dh)
    check(azb_memcpy_h2d(dst.raw(), src.raw(), bytes), "azb_memcpy_h2d");
  else if (sh)
    check(azb_memcpy_d2h(dst.raw(), src.raw(), bytes), "azb_memcpy_d2h");
  else
    std::copy(src.raw(), src.raw() + dst.size(), dst.raw());
}

template <int Dim>
struct Grid {
  static constexpr int dim_v = Dim;
  long N_phys = 0, N_phys_pad = 0, N_fourier = 0, N_fourier_pad = 0;
  Grid() = default;
  explicit Grid(long N) : Grid(N, 3 * N / 2) {}
  Grid(long N, long M) : N_phys(N), N_phys_pad(M), N_fourier(N / 2 + 1), N_fourier_pad(M / 2 + 1) {}
  shape_t<Dim + 1> shape_fourier(long n_vars) const { return shape(n_vars, N_phys, N_fourier); }
  shape_t<Dim + 1> shape_fourier_pad(long n_vars) const { return shape(n_vars, N_phys_pad, N_fourier_pad); }
  shape_t<Dim + 1> shape_phys(long n_vars) const { return shape(n_vars, N_phys, N_phys); }
  shape_t<Dim + 1> shape_phys_pad(long n_vars) const { return shape(n_vars, N_phys_pad, N_phys_pad); }

 private:
  static shape_t<Dim + 1> shape(long n_vars, long full, long last) {
    shape_t<Dim + 1> s;
    s[0] = n_vars;
    for (int i = 1; i < Dim; ++i) s[i] = full;
    s[Dim] = last;
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

// ---------------------------------------------------------------------------
// Equation<Dim> (equations/equation.hpp:27-56)
// ---------------------------------------------------------------------------
template <int Dim>
class Equation {
 public:
  static constexpr int dim_v = Dim;
  Equation(const Grid<Dim>& grid) : grid_(grid) {}
  virtual ~Equation() = default;
  // replaces dudt with the time derivative of u; dudt may alias u
  virtual void dudt(const array_view<complex_t, Dim + 1>& dudt, const array_const_view<complex_t, Dim + 1>& u,
                    double t, double dt, double C) = 0;
  virtual double dt(double C) const = 0;
  virtual int n_vars() const = 0;
  virtual real_t visc() const { return 0; }
  virtual void* get_fft_work_area() const { return nullptr; }
  const Grid<Dim>& grid() const { return grid_; }

 protected:
  Grid<Dim> grid_;
};

enum class precision { fp64 = AZB_FP64, fp32_mixed = AZB_FP32_MIXED };

// IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> on the MI355X.  The solver
// handle owns the workspace the reference's constructor allocates
// (incompressible_euler.hpp:56-96) and the RK3 stage buffers.  Like the reference's
// single-device IncompressibleEuler, it shares no FFT work area
// (get_fft_work_area() is the base's nullptr): its intermediates are in use during
// every dudt.
template <int Dim>
class IncompressibleEuler final : public Equation<Dim> {
 public:
  IncompressibleEuler(const Grid<Dim>& grid, const SmoothCutoff1D& visc, device_type device = device_type::hip,
                      precision prec = precision::fp64, void* stream = nullptr)
      : Equation<Dim>(grid), visc_(visc) {
    if (device != device_type::hip) fatal("IncompressibleEuler on the MI355X needs device_type::hip");
    azb_solver* h = nullptr;
    check(azb_create(Dim, (int)grid.N_phys, (int)grid.N_phys_pad, Dim, static_cast<int>(prec),
                     azb_sv_params{visc.eps, visc.s, visc.k0}, stream, &h),
          "azb_create");
    handle_.reset(h);
  }
  void dudt(const array_view<complex_t, Dim + 1>& dudt_hat, const array_const_view<complex_t, Dim + 1>& u_hat,
            double, double, double) override {
    if (dudt_hat.shape(0) != n_vars() || u_hat.shape(0) != n_vars()) fatal("Wrong number of variables");
    check_device(dudt_hat.memory_location(), u_hat.memory_location());
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
  static void check_device(device_type a, device_type b) {
    if (a != device_type::hip || b != device_type::hip) fatal("IncompressibleEuler::dudt needs device arrays");
  }
  struct Del {
    void operator()(azb_solver* s) const { azb_destroy(s); }
  };
  SmoothCutoff1D visc_;
  std::unique_ptr<azb_solver, Del> handle_;
};

// y <- a x + b y on device arrays (operations/axpby.hpp:28-56)
template <int Rank>
void axpby(double a, const array_const_view<complex_t, Rank>& x, double b, const array_view<complex_t, Rank>& y) {
  if (x.size() != y.size()) fatal("axpby: size mismatch");
  check(azb_axpby(a, x.raw(), b, y.raw(), y.size(), nullptr), "azb_axpby");
}

// ---------------------------------------------------------------------------
// TimeIntegrator<Dim> / SSP_RK3<Dim> (evolution/time_integrator.hpp:30-60,
// evolution/ssp_rk3.hpp:28-64)
// ---------------------------------------------------------------------------
template <int Dim>
class TimeIntegrator {
 public:
  static constexpr int dim_v = Dim;
  TimeIntegrator() = delete;
  TimeIntegrator(device_type device, const std::shared_ptr<Equation<Dim>>& equation)
      : device_(device), equation_(equation) {}
  virtual ~TimeIntegrator() = default;
  virtual double integrate(double t, double max_dt, double C, const array_view<complex_t, Dim + 1>& u) = 0;
  device_type memory_location() const { return device_; }
  std::shared_ptr<Equation<Dim>> equation() { return equation_; }
  std::shared_ptr<const Equation<Dim>> equation() const { return equation_; }

 protected:
  device_type device_;
  std::shared_ptr<Equation<Dim>> equation_;
};

// Shu-Osher SSP-RK3.  With an IncompressibleEuler equation the three stages and five
// axpbys run fused on the device (azb_rk3_step: dt from stage 1's CFL stays on the
// device); any other Equation takes the reference's sequence of dudt calls and
// device axpbys on two stage arrays of `shape` (allocated on first use).
template <int Dim>
class SSP_RK3 final : public TimeIntegrator<Dim> {
  using super = TimeIntegrator<Dim>;

 public:
  SSP_RK3() = delete;
  SSP_RK3(device_type device, const shape_t<Dim + 1>& shape, const std::shared_ptr<Equation<Dim>>& equation)
      : super(device, equation), shape_(shape) {
    if (device != device_type::hip) fatal("SSP_RK3 on the MI355X needs device_type::hip");
  }
  SSP_RK3(const SSP_RK3&) = delete;
  SSP_RK3& operator=(const SSP_RK3&) = delete;

  double integrate(double t, double max_dt, double C, const array_view<complex_t, Dim + 1>& u) override {
    if (auto* ie = dynamic_cast<IncompressibleEuler<Dim>*>(this->equation_.get())) {
      double dt = 0;
      check(azb_rk3_step(ie->handle(), u.raw(), max_dt, C, &dt), "azb_rk3_step");
      return dt;
    }
    if (!u1_) {
      u1_ = std::make_unique<device_array<complex_t, Dim + 1>>(shape_);
      u2_ = std::make_unique<device_array<complex_t, Dim + 1>>(shape_);
    }
    Equation<Dim>& eq = *this->equation_;
    const array_view<complex_t, Dim + 1> u1 = u1_->view(), u2 = u2_->view();
    eq.dudt(u1, u, t, max_dt, C);
    const double dt = std::min(eq.dt(C), max_dt);
    axpby<Dim + 1>(1, u, dt, u1);
    eq.dudt(u2, u1, t, dt, C);
    axpby<Dim + 1>(1, u1, dt, u2);
    axpby<Dim + 1>(3. / 4, u, 1. / 4, u2);
    eq.dudt(u1, u2, t, dt, C);
    axpby<Dim + 1>(1, u2, dt, u1);
    axpby<Dim + 1>(2. / 3, u1, 1. / 3, u);
    check(azb_device_synchronize(), "azb_device_synchronize");
    return dt;
  }

 private:
  shape_t<Dim + 1> shape_;
  std::unique_ptr<device_array<complex_t, Dim + 1>> u1_, u2_;
};

// ---------------------------------------------------------------------------
// Simulation<Dim> (simulation.hpp:30-93, simulation.cpp:6-102)
// ---------------------------------------------------------------------------
template <int Dim>
class Simulation {
 public:
  static constexpr int dim_v = Dim;
  Simulation() = delete;
  Simulation(const Grid<Dim>& grid, double C, const std::shared_ptr<TimeIntegrator<Dim>>& timestepper,
             device_type device = device_type::hip)
      : u_(alloc(grid.shape_fourier(timestepper->equation()->n_vars()), device)),
        C_(C), grid_(grid), timestepper_(timestepper), memory_location_(device) {}
  Simulation(const array_const_view<complex_t, Dim + 1>& u, const Grid<Dim>& grid, double C,
             const std::shared_ptr<TimeIntegrator<Dim>>& timestepper)
      : u_(alloc(u.shape(), device_type::hip)), C_(C), grid_(grid), timestepper_(timestepper),
        memory_location_(device_type::hip) {
    azeban::copy<complex_t, Dim + 1>(u_->view(), u);
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;
  Simulation(Simulation&&) = default;
  Simulation& operator=(Simulation&&) = default;

  // host-driven loop, exactly Simulation::simulate_until (simulation.cpp:89-102)
  void simulate_until(double t) {
    const double pi = 3.14159265358979323846;
    while (time_ < t) {
      const real_t eps = equation()->visc();
      const double max_dt = std::min(t - time_, C_ * 2. / (eps * (pi * grid_.N_phys) * (pi * grid_.N_phys)));
      const double dt = timestepper_->integrate(t, max_dt, C_, u());
      if (dt <= 1e-10) std::fprintf(stderr, "Warning: Timestep is tiny. dt = %g\n", dt);
      time_ += dt;
      ++steps_;
    }
    time_ = t;
  }
  void simulate_for(double t) { simulate_until(time_ + t); }
  // device-driven loop of an IncompressibleEuler solve: dt and the time stay on the
  // device, the host checks completion every `chunk` steps (same steps as simulate_until)
  void simulate_until_device(double t, int chunk = 8) {
    auto* ie = dynamic_cast<IncompressibleEuler<Dim>*>(equation().get());
    if (!ie) fatal("simulate_until_device needs an IncompressibleEuler equation");
    int steps = 0;
    check(azb_simulate_until(ie->handle(), u_->raw(), t, C_, &time_, 0, chunk, &steps), "azb_simulate_until");
    steps_ += steps;
  }

  void reset() { time_ = 0; }
  void set_time(double t) { time_ = t; }
  double time() const { return time_; }
  array_view<complex_t, Dim + 1> u() { return u_->view(); }
  array_const_view<complex_t, Dim + 1> u() const { return u_->view(); }
  const Grid<Dim>& grid() const { return grid_; }
  long n_vars() const { return u_->shape(0); }
  device_type memory_location() const { return memory_location_; }
  std::shared_ptr<Equation<Dim>> equation() { return timestepper_->equation(); }
  std::shared_ptr<const Equation<Dim>> equation() const {
    const std::shared_ptr<const TimeIntegrator<Dim>> ts = timestepper_;
    return ts->equation();
  }
  long steps() const { return steps_; }

 private:
  static std::unique_ptr<device_array<complex_t, Dim + 1>> alloc(const shape_t<Dim + 1>& s, device_type d) {
    if (d != device_type::hip) fatal("Simulation on the MI355X keeps u in device memory (device_type::hip)");
    return std::make_unique<device_array<complex_t, Dim + 1>>(s);
  }
  std::unique_ptr<device_array<complex_t, Dim + 1>> u_;
  double C_;
  Grid<Dim> grid_;
  std::shared_ptr<TimeIntegrator<Dim>> timestepper_;
  double time_ = 0;
  device_type memory_location_;
  long steps_ = 0;
};

// ---------------------------------------------------------------------------
// FFT<Dim, ScalarU> (operations/fft.hpp:37-121) and make_fft (fft_factory.hpp:27-60)
// ---------------------------------------------------------------------------
enum fft_direction { FFT_FORWARD = 1, FFT_BACKWARD = 2 };

template <int Dim, typename ScalarU = real_t>
class FFT {
 public:
  static constexpr int dim_v = Dim;
  using scalar_u_t = ScalarU;
  static_assert(std::is_same_v<scalar_u_t, real_t> || std::is_same_v<scalar_u_t, complex_t>,
                "ScalarU must be eiter real_t or complex_t");

  template <bool enable = Dim == 1, typename = std::enable_if_t<enable>>
  FFT(int direction, bool transform_x = true) : direction_(direction), transform_dims_{transform_x} {}
  template <bool enable = Dim == 2, typename = std::enable_if_t<enable>>
  FFT(int direction, bool transform_x = true, bool transform_y = true)
      : direction_(direction), transform_dims_{transform_x, transform_y} {}
  template <bool enable = Dim == 3, typename = std::enable_if_t<enable>>
  FFT(int direction, bool transform_x = true, bool transform_y = true, bool transform_z = true)
      : direction_(direction), transform_dims_{transform_x, transform_y, transform_z} {}
  virtual ~FFT() = default;

  // binds the arrays; with allocate_work_area the transform owns its work area,
  // otherwise set_work_area must provide get_work_area_size() bytes before use
  void initialize(const array_view<complex_t, Dim + 1>& u_hat, const array_view<scalar_u_t, Dim + 1>& u,
                  bool allocate_work_area = true) {
    u_hat_ = u_hat;
    u_ = u;
    data_dim_ = (int)u.shape(0);
    do_initialize(u_hat, u, allocate_work_area);
  }
  virtual size_t get_work_area_size() const = 0;
  void set_work_area(void* work_area) { do_set_work_area(work_area); }

  virtual void forward() = 0;
  virtual void backward() = 0;

  bool is_forward() const { return direction_ & FFT_FORWARD; }
  bool is_backward() const { return direction_ & FFT_BACKWARD; }

  const shape_t<Dim + 1>& shape() const { return u_.shape(); }
  long shape(int i) const { return u_.shape(i); }

  const array_view<complex_t, Dim + 1>& u_hat() { return u_hat_; }
  array_const_view<complex_t, Dim + 1> u_hat() const { return u_hat_; }
  const array_view<scalar_u_t, Dim + 1>& u() { return u_; }
  array_const_view<scalar_u_t, Dim + 1> u() const { return u_; }

  // real transforms keep n/2+1 entries of the last transformed axis
  shape_t<Dim + 1> shape_u_hat(const shape_t<Dim + 1>& shape_u) const {
    shape_t<Dim + 1> s = shape_u;
    if (std::is_same_v<scalar_u_t, real_t>) s[Dim] = shape_u[Dim] / 2 + 1;
    return s;
  }
  shape_t<Dim + 1> shape_u(const shape_t<Dim + 1>& shape_u_hat) const {
    shape_t<Dim + 1> s = shape_u_hat;
    if (std::is_same_v<scalar_u_t, real_t>) s[Dim] = 2 * (shape_u_hat[Dim] - 1);
    return s;
  }

  virtual void* get_work_area() const { return nullptr; }

 protected:
  int direction_ = FFT_FORWARD | FFT_BACKWARD;
  bool transform_dims_[Dim];
  array_view<complex_t, Dim + 1> u_hat_;
  array_view<scalar_u_t, Dim + 1> u_;
  int data_dim_ = 0;

  virtual void do_initialize(const array_view<complex_t, Dim + 1>& u_hat, const array_view<scalar_u_t, Dim + 1>& u,
                             bool allocate_work_area) = 0;
  virtual void do_set_work_area(void* work_area) = 0;
};

// Real-to-complex transform of every axis on the MI355X (the CUFFT_R2C role,
// src/azeban/cuda/operations/cufft.cpp:18-262): forward = unnormalised r2c with
// e^{-ikx}, backward = c2r with e^{+ikx}, which may overwrite u_hat as FFTW's and
// cuFFT's do; batched over the leading (variable) axis.  Every transformed axis must
// have the same built length.  The work area holds the twiddle table.
template <int Dim>
class HIPFFT_R2C final : public FFT<Dim, real_t> {
  using super = FFT<Dim, real_t>;

 public:
  using super::super;
  ~HIPFFT_R2C() override {
    if (owned_) azb_free(owned_);
  }
  size_t get_work_area_size() const override {
    const size_t b = azb_fft_work_bytes(n_, AZB_FP64);
    if (b == 0) fatal("FFT length not built");
    return b;
  }
  void forward() override {
    if (!this->is_forward()) fatal("FFT was not planned for the forward direction");
    check(azb_fft_r2c_ws(Dim, n_, (int)this->u_.shape(0), AZB_FP64, this->u_.raw(), this->u_hat_.raw(), work_,
                         nullptr),
          "azb_fft_r2c_ws");
    check(azb_device_synchronize(), "azb_device_synchronize");
  }
  void backward() override {
    if (!this->is_backward()) fatal("FFT was not planned for the backward direction");
    check(azb_fft_c2r_ws(Dim, n_, (int)this->u_.shape(0), AZB_FP64, this->u_hat_.raw(), this->u_.raw(), work_,
                         nullptr),
          "azb_fft_c2r_ws");
    check(azb_device_synchronize(), "azb_device_synchronize");
  }
  void* get_work_area() const override { return work_; }

 protected:
  void do_initialize(const array_view<complex_t, Dim + 1>& u_hat, const array_view<real_t, Dim + 1>& u,
                     bool allocate_work_area) override {
    for (int d = 0; d < Dim; ++d)
      if (!this->transform_dims_[d]) fatal("HIPFFT_R2C transforms every axis");
    if (u_hat.memory_location() != device_type::hip || u.memory_location() != device_type::hip)
      fatal("HIPFFT_R2C needs device arrays");
    n_ = (int)u.shape(1);
    for (int d = 1; d <= Dim; ++d)
      if (u.shape(d) != n_) fatal("HIPFFT_R2C needs the same length on every axis");
    if (u_hat.shape() != this->shape_u_hat(u.shape())) fatal("u_hat does not have the r2c shape of u");
    if (allocate_work_area) {
      void* p = nullptr;
      check(azb_malloc(&p, get_work_area_size()), "azb_malloc");
      owned_ = p;
      do_set_work_area(p);
    }
  }
  void do_set_work_area(void* work_area) override {
    check(azb_fft_work_init(work_area, n_, AZB_FP64), "azb_fft_work_init");
    work_ = work_area;
  }

 private:
  int n_ = 0;
  void* work_ = nullptr;
  void* owned_ = nullptr;
};

template <int Dim, typename ScalarU = real_t, typename... T>
std::shared_ptr<FFT<Dim, ScalarU>> make_fft(device_type device, int direction = FFT_FORWARD | FFT_BACKWARD,
                                            T... transform) {
  static_assert(sizeof...(T) == Dim || sizeof...(T) == 0, "one transform flag per axis");
  static_assert(std::is_same_v<ScalarU, real_t>, "only real-to-complex transforms are built");
  if (device != device_type::hip) fatal("Unsupported combination of memory loctions");
  return std::make_shared<HIPFFT_R2C<Dim>>(direction, transform...);
}

template <int Dim>
std::shared_ptr<FFT<Dim, real_t>> make_fft(const array_view<complex_t, Dim + 1>& u_hat,
                                           const array_view<real_t, Dim + 1>& u,
                                           int direction = FFT_FORWARD | FFT_BACKWARD) {
  if (u_hat.memory_location() != u.memory_location()) fatal("Memory locations mismatch");
  auto fft = make_fft<Dim, real_t>(u.memory_location(), direction);
  fft->initialize(u_hat, u);
  return fft;
}

// ---------------------------------------------------------------------------
// Leray projection and the energy spectrum on device arrays
// ---------------------------------------------------------------------------
template <int Dim>
void leray(const array_view<complex_t, Dim + 1>& u_hat) {
  check(azb_leray(Dim, (int)u_hat.shape(1), u_hat.raw(), nullptr), "azb_leray");
}

template <int Dim>
std::vector<real_t> energy_spectrum(const Grid<Dim>& grid, const array_const_view<complex_t, Dim + 1>& u_hat) {
  std::vector<real_t> e(grid.N_fourier);
  check(azb_energy_spectrum(Dim, (int)grid.N_phys, u_hat.raw(), e.data(), nullptr), "azb_energy_spectrum");
  return e;
}

}  // namespace azeban

#endif  // AZEBAN_HIP_HPP_
