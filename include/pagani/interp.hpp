// interp.hpp -- interpolation tables for stateful integrands (device side).
//
// Replaces quad::Interp1D / Interp2D / Interp3D (common/cuda/Interp1D.cuh,
// Interp2D.cuh, Interp3D.cuh): knot vectors and a value table in device
// memory, evaluated inside integrands by linear, bilinear and trilinear
// interpolation.  Same class names, constructors, table layouts, search
// steps and formulas (operation order kept, FMA contraction off), so a
// reference integrand that holds these members compiles unchanged against
// this header and gets the same values.
//
// Differences, all outside the reference's defined behaviour:
// - Copies share one immutable device allocation (reference counted on the
//   host) instead of deep-copying it.  An integrand is copied into every
//   evaluation launch; sharing makes that free and leak-free.
// - Values outside the knot range use the end cell (linear extrapolation);
//   the reference's searches loop forever (2D, 3D) or divide by zero (1D)
//   there.  clamp() evaluates at the clamped point, as in the reference.
// - Tables live in plain device memory (the reference's stream reader uses
//   managed memory).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <istream>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pagani/str_to_doubles.hh"

namespace pagani {
namespace detail {

// One device allocation shared by every copy of an interpolator.
struct TableOwner {
  std::atomic<long> refs{1};
  double* dev = nullptr;
};

inline TableOwner* upload_tables(const std::vector<std::pair<const double*, std::size_t>>& parts,
                                 std::vector<double*>& where)
{
  std::size_t total = 0;
  for (const auto& p : parts) total += p.second;
  TableOwner* own = new TableOwner();
  if (hipMalloc(&own->dev, total * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    delete own;
    throw std::bad_alloc();
  }
  std::vector<double> host;
  host.reserve(total);
  where.clear();
  for (const auto& p : parts) {
    where.push_back(own->dev + host.size());
    host.insert(host.end(), p.first, p.first + p.second);
  }
  if (hipMemcpy(own->dev, host.data(), total * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(own->dev);
    delete own;
    throw std::runtime_error("interpolation table upload failed");
  }
  return own;
}

inline void retain(TableOwner* own)
{
  if (own) own->refs.fetch_add(1, std::memory_order_relaxed);
}

inline void release(TableOwner*& own)
{
  if (own && own->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
    if (own->dev) (void)hipFree(own->dev);
    delete own;
  }
  own = nullptr;
}

// The reference's searches (Interp1D::_find_smallest__index_range,
// Interp1D.cuh:186-202; FindNeighbourIndices, Interp2D.cuh:209-233) for
// values inside the table; outside it (or NaN) the end cell.  The loops
// take at most log2(n) + 2 steps for values inside; the step bound only
// makes termination evident.
__host__ __device__ inline void cell_1d(const double* xs, std::size_t n, double val,
                                        std::size_t& a, std::size_t& b)
{
  a = 0;
  b = 1;
  if (!(xs[0] <= val)) return;
  if (!(val <= xs[n - 1])) {
    a = n - 2;
    b = n - 1;
    return;
  }
  std::size_t left = 0, right = n - 1;
  for (int step = 0; step < 128 && left < right; ++step) {
    const std::size_t mid = static_cast<std::size_t>((left + right) * 0.5);
    if (xs[mid] <= val && xs[mid + 1] >= val) {
      a = mid;
      b = mid + 1;
      return;
    }
    if (xs[mid] > val)
      right = mid;
    else
      left = mid + 1;
  }
  a = left;
  b = right;
}

__host__ __device__ inline void neighbours(const double* arr, std::size_t n, double val,
                                           std::size_t& leftI, std::size_t& rightI)
{
  leftI = 0;
  rightI = 1;
  if (!(arr[0] <= val)) return;
  if (!(val <= arr[n - 1])) {
    leftI = n - 2;
    rightI = n - 1;
    return;
  }
  std::size_t lo = 0, hi = n - 1;
  for (int step = 0; step < 128 && lo <= hi; ++step) {
    const std::size_t mid = static_cast<std::size_t>((hi + lo) * 0.5);
    if (arr[mid] <= val && arr[mid + 1] >= val) {
      leftI = mid;
      rightI = mid + 1;
      return;
    }
    if (arr[mid] > val)
      hi = mid;
    else
      lo = mid;
  }
}

__host__ __device__ inline double do_clamp(double v, double lo, double hi)
{
  return (v < lo) ? lo : (hi < v) ? hi : v;
}

inline void check_knots(std::size_t n, const char* what)
{
  if (n < 2) throw std::invalid_argument(std::string(what) + ": at least two knots per axis");
}

}  // namespace detail
}  // namespace pagani

namespace quad {

// Interp1D (common/cuda/Interp1D.cuh:28-63): z(x) on knots xs.
class Interp1D {
 public:
  Interp1D() = default;
  Interp1D(const double* xs, const double* zs, std::size_t cols) { init(xs, zs, cols); }
  template <std::size_t M>
  Interp1D(std::array<double, M> const& xs, std::array<double, M> const& zs)
  {
    init(xs.data(), zs.data(), M);
  }
  Interp1D(std::vector<double> const& xs, std::vector<double> const& zs)
  {
    if (xs.size() != zs.size()) throw std::invalid_argument("Interp1D: size mismatch");
    init(xs.data(), zs.data(), xs.size());
  }
  Interp1D(const Interp1D& o) : _cols(o._cols), _xs(o._xs), _zs(o._zs), _own(o._own)
  {
    pagani::detail::retain(_own);
  }
  Interp1D& operator=(Interp1D const& rhs)
  {
    Interp1D tmp(rhs);
    swap(tmp);
    return *this;
  }
  ~Interp1D() { pagani::detail::release(_own); }

  void swap(Interp1D& o)
  {
    std::swap(_cols, o._cols);
    std::swap(_xs, o._xs);
    std::swap(_zs, o._zs);
    std::swap(_own, o._own);
  }

  std::size_t get_device_mem_footprint() const { return 8 * 2 * _cols; }

  __device__ double operator()(double x) const
  {
#pragma clang fp contract(off)
    std::size_t i0, i1;
    pagani::detail::cell_1d(_xs, _cols, x, i0, i1);
    const double y0 = _zs[i0];
    const double y1 = _zs[i1];
    const double x0 = _xs[i0];
    const double x1 = _xs[i1];
    return (y0 * (x1 - x) + y1 * (x - x0)) / (x1 - x0);
  }
  __device__ double eval(double x) const { return (*this)(x); }
  __device__ double min_x() const { return _xs[0]; }
  __device__ double max_x() const { return _xs[_cols - 1]; }
  __device__ double do_clamp(double v, double lo, double hi) const
  {
    return pagani::detail::do_clamp(v, lo, hi);
  }
  __device__ double clamp(double x) const { return eval(do_clamp(x, min_x(), max_x())); }

  // two lines: knots, values (operator>>, Interp1D.cuh:226-249)
  friend std::istream& operator>>(std::istream& is, Interp1D& interp)
  {
    std::string buffer;
    std::getline(is, buffer);
    const std::vector<double> xs = str_to_doubles(buffer);
    std::getline(is, buffer);
    const std::vector<double> zs = str_to_doubles(buffer);
    Interp1D tmp(xs, zs);
    interp.swap(tmp);
    return is;
  }

 private:
  void init(const double* xs, const double* zs, std::size_t cols)
  {
    pagani::detail::check_knots(cols, "Interp1D");
    std::vector<double*> at;
    _own = pagani::detail::upload_tables({{xs, cols}, {zs, cols}}, at);
    _cols = cols;
    _xs = at[0];
    _zs = at[1];
  }

  std::size_t _cols = 0;
  const double* _xs = nullptr;
  const double* _zs = nullptr;
  pagani::detail::TableOwner* _own = nullptr;
};

// Interp2D (common/cuda/Interp2D.cuh:15-298): z(x, y) with the table
// zs[row * cols + col], rows along ys, columns along xs.
class Interp2D {
 public:
  Interp2D() = default;
  Interp2D(const double* xs, const double* ys, const double* zs, std::size_t cols,
           std::size_t rows)
  {
    init(xs, ys, zs, cols, rows);
  }
  template <std::size_t M, std::size_t N>
  Interp2D(std::array<double, M> const& xs, std::array<double, N> const& ys,
           std::array<double, N * M> const& zs)
  {
    init(xs.data(), ys.data(), zs.data(), M, N);
  }
  Interp2D(std::vector<double> const& xs, std::vector<double> const& ys,
           std::vector<double> const& zs)
  {
    if (zs.size() != xs.size() * ys.size()) throw std::invalid_argument("Interp2D: size mismatch");
    init(xs.data(), ys.data(), zs.data(), xs.size(), ys.size());
  }
  // values given as zs[col][row] (Interp2D.cuh:136-156)
  template <std::size_t M, std::size_t N>
  Interp2D(std::array<double, M> xs, std::array<double, N> ys,
           std::array<std::array<double, N>, M> zs)
  {
    std::vector<double> buffer(N * M);
    for (std::size_t i = 0; i < M; ++i)
      for (std::size_t j = 0; j < N; ++j) buffer[i + j * M] = zs[i][j];
    init(xs.data(), ys.data(), buffer.data(), M, N);
  }
  Interp2D(const Interp2D& o)
      : _rows(o._rows), _cols(o._cols), interpT(o.interpT), interpR(o.interpR),
        interpC(o.interpC), _own(o._own)
  {
    pagani::detail::retain(_own);
  }
  Interp2D& operator=(Interp2D const& rhs)
  {
    Interp2D tmp(rhs);
    swap(tmp);
    return *this;
  }
  ~Interp2D() { pagani::detail::release(_own); }

  void swap(Interp2D& o)
  {
    std::swap(_rows, o._rows);
    std::swap(_cols, o._cols);
    std::swap(interpT, o.interpT);
    std::swap(interpR, o.interpR);
    std::swap(interpC, o.interpC);
    std::swap(_own, o._own);
  }

  std::size_t get_device_mem_footprint() const { return 8 * (_cols * _rows + _cols + _rows); }

  __device__ double operator()(double x, double y) const
  {
#pragma clang fp contract(off)
    std::size_t y1 = 0, y2 = 0, x1 = 0, x2 = 0;
    pagani::detail::neighbours(interpR, _rows, y, y1, y2);
    pagani::detail::neighbours(interpC, _cols, x, x1, x2);
    const double q11 = interpT[y1 * _cols + x1];
    const double q12 = interpT[y2 * _cols + x1];
    const double q21 = interpT[y1 * _cols + x2];
    const double q22 = interpT[y2 * _cols + x2];
    const double x1_val = interpC[x1];
    const double x2_val = interpC[x2];
    const double y1_val = interpR[y1];
    const double y2_val = interpR[y2];
    const double f_x_y1 = q11 * (x2_val - x) / (x2_val - x1_val) +
                          q21 * (x - x1_val) / (x2_val - x1_val);
    const double f_x_y2 = q12 * (x2_val - x) / (x2_val - x1_val) +
                          q22 * (x - x1_val) / (x2_val - x1_val);
    return f_x_y1 * (y2_val - y) / (y2_val - y1_val) + f_x_y2 * (y - y1_val) / (y2_val - y1_val);
  }
  __device__ double eval(double x, double y) const { return (*this)(x, y); }
  __device__ double min_x() const { return interpC[0]; }
  __device__ double max_x() const { return interpC[_cols - 1]; }
  __device__ double min_y() const { return interpR[0]; }
  __device__ double max_y() const { return interpR[_rows - 1]; }
  __device__ double do_clamp(double v, double lo, double hi) const
  {
    return pagani::detail::do_clamp(v, lo, hi);
  }
  __device__ double clamp(double x, double y) const
  {
    return eval(do_clamp(x, min_x(), max_x()), do_clamp(y, min_y(), max_y()));
  }

  // three lines: x knots, y knots, values (operator>>, Interp2D.cuh:183-207)
  friend std::istream& operator>>(std::istream& is, Interp2D& interp)
  {
    std::string buffer;
    std::getline(is, buffer);
    const std::vector<double> xs = str_to_doubles(buffer);
    std::getline(is, buffer);
    const std::vector<double> ys = str_to_doubles(buffer);
    std::getline(is, buffer);
    const std::vector<double> zs = str_to_doubles(buffer);
    Interp2D tmp(xs, ys, zs);
    interp.swap(tmp);
    return is;
  }

 private:
  void init(const double* xs, const double* ys, const double* zs, std::size_t cols,
            std::size_t rows)
  {
    pagani::detail::check_knots(cols, "Interp2D");
    pagani::detail::check_knots(rows, "Interp2D");
    std::vector<double*> at;
    _own = pagani::detail::upload_tables({{ys, rows}, {xs, cols}, {zs, rows * cols}}, at);
    _rows = rows;
    _cols = cols;
    interpR = at[0];
    interpC = at[1];
    interpT = at[2];
  }

  std::size_t _rows = 0;
  std::size_t _cols = 0;
  const double* interpT = nullptr;
  const double* interpR = nullptr;
  const double* interpC = nullptr;
  pagani::detail::TableOwner* _own = nullptr;
};

// Interp3D (common/cuda/Interp3D.cuh:16-298): v(x, y, z) with the table
// vs[x + nx*y + nx*ny*z].
class Interp3D {
 public:
  Interp3D() = default;
  Interp3D(const double* xs, const double* ys, const double* zs, const double* vs,
           std::size_t x_size, std::size_t y_size, std::size_t z_size)
  {
    init(xs, ys, zs, vs, x_size, y_size, z_size);
  }
  template <std::size_t M, std::size_t N, std::size_t S>
  Interp3D(std::array<double, M> const& xs, std::array<double, N> const& ys,
           std::array<double, S> const& zs, std::array<double, M * N * S> const& vs)
  {
    init(xs.data(), ys.data(), zs.data(), vs.data(), M, N, S);
  }
  Interp3D(std::vector<double> const& xs, std::vector<double> const& ys,
           std::vector<double> const& zs, std::vector<double> const& vs)
  {
    if (vs.size() != xs.size() * ys.size() * zs.size())
      throw std::invalid_argument("Interp3D: size mismatch");
    init(xs.data(), ys.data(), zs.data(), vs.data(), xs.size(), ys.size(), zs.size());
  }
  Interp3D(const Interp3D& o)
      : size_x(o.size_x), size_y(o.size_y), size_z(o.size_z), _xs(o._xs), _ys(o._ys),
        _zs(o._zs), interpT(o.interpT), _own(o._own)
  {
    pagani::detail::retain(_own);
  }
  Interp3D& operator=(Interp3D const& rhs)
  {
    Interp3D tmp(rhs);
    swap(tmp);
    return *this;
  }
  ~Interp3D() { pagani::detail::release(_own); }

  void swap(Interp3D& o)
  {
    std::swap(size_x, o.size_x);
    std::swap(size_y, o.size_y);
    std::swap(size_z, o.size_z);
    std::swap(_xs, o._xs);
    std::swap(_ys, o._ys);
    std::swap(_zs, o._zs);
    std::swap(interpT, o.interpT);
    std::swap(_own, o._own);
  }

  std::size_t get_device_mem_footprint() const
  {
    return 8 * (size_x * size_y * size_z + size_x + size_y + size_z);
  }

  __device__ std::size_t index(std::size_t x, std::size_t y, std::size_t z) const
  {
    return x + (size_x * y) + (size_x * size_y * z);
  }

  __device__ double operator()(double x, double y, double z) const
  {
#pragma clang fp contract(off)
    std::size_t y0 = 0, y1 = 0, x0 = 0, x1 = 0, z0 = 0, z1 = 0;
    pagani::detail::neighbours(_ys, size_y, y, y0, y1);
    pagani::detail::neighbours(_xs, size_x, x, x0, x1);
    pagani::detail::neighbours(_zs, size_z, z, z0, z1);
    const double x_d = (x - _xs[x0]) / (_xs[x1] - _xs[x0]);
    const double y_d = (y - _ys[y0]) / (_ys[y1] - _ys[y0]);
    const double z_d = (z - _zs[z0]) / (_zs[z1] - _zs[z0]);
    const double c00 = interpT[index(x0, y0, z0)] * (1 - x_d) + interpT[index(x1, y0, z0)] * x_d;
    const double c01 = interpT[index(x0, y0, z1)] * (1 - x_d) + interpT[index(x1, y0, z1)] * x_d;
    const double c10 = interpT[index(x0, y1, z0)] * (1 - x_d) + interpT[index(x1, y1, z0)] * x_d;
    const double c11 = interpT[index(x0, y1, z1)] * (1 - x_d) + interpT[index(x1, y1, z1)] * x_d;
    const double c0 = c00 * (1 - y_d) + c10 * y_d;
    const double c1 = c01 * (1 - y_d) + c11 * y_d;
    return c0 * (1 - z_d) + c1 * z_d;
  }
  __device__ double eval(double x, double y, double z) const { return (*this)(x, y, z); }
  __device__ double min_x() const { return _xs[0]; }
  __device__ double max_x() const { return _xs[size_x - 1]; }
  __device__ double min_y() const { return _ys[0]; }
  __device__ double max_y() const { return _ys[size_y - 1]; }
  __device__ double min_z() const { return _zs[0]; }
  __device__ double max_z() const { return _zs[size_z - 1]; }
  __device__ double do_clamp(double v, double lo, double hi) const
  {
    return pagani::detail::do_clamp(v, lo, hi);
  }
  __device__ double clamp(double x, double y, double z) const
  {
    return eval(do_clamp(x, min_x(), max_x()), do_clamp(y, min_y(), max_y()),
                do_clamp(z, min_z(), max_z()));
  }

 private:
  void init(const double* xs, const double* ys, const double* zs, const double* vs,
            std::size_t nx, std::size_t ny, std::size_t nz)
  {
    pagani::detail::check_knots(nx, "Interp3D");
    pagani::detail::check_knots(ny, "Interp3D");
    pagani::detail::check_knots(nz, "Interp3D");
    std::vector<double*> at;
    _own = pagani::detail::upload_tables({{xs, nx}, {ys, ny}, {zs, nz}, {vs, nx * ny * nz}}, at);
    size_x = nx;
    size_y = ny;
    size_z = nz;
    _xs = at[0];
    _ys = at[1];
    _zs = at[2];
    interpT = at[3];
  }

  std::size_t size_x = 0, size_y = 0, size_z = 0;
  const double* _xs = nullptr;
  const double* _ys = nullptr;
  const double* _zs = nullptr;
  const double* interpT = nullptr;
  pagani::detail::TableOwner* _own = nullptr;
};

}  // namespace quad
