// volume.hh -- quad::Volume<T, NDIM>, the integration box (common/cuda/
// Volume.cuh:13-52 of the reference: defaults to the unit cube, same
// constructors).
#pragma once

#include <array>
#include <cstddef>
#include <cstring>

namespace quad {

template <typename T, int NDIM>
struct Volume {
  T lows[NDIM] = {0.0};
  T highs[NDIM];

  Volume()
  {
    for (T& x : highs) x = 1.0;
  }

  Volume(std::array<T, NDIM> l, std::array<T, NDIM> h)
  {
    std::memcpy(lows, l.data(), NDIM * sizeof(T));
    std::memcpy(highs, h.data(), NDIM * sizeof(T));
  }

  Volume(T const* l, T const* h)
  {
    std::memcpy(lows, l, NDIM * sizeof(T));
    std::memcpy(highs, h, NDIM * sizeof(T));
  }

  Volume(T l, T h)
  {
    for (int i = 0; i < NDIM; ++i) {
      lows[i] = l;
      highs[i] = h;
    }
  }
};

}  // namespace quad
