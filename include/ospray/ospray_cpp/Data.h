// OSPRay C++ wrapper: data arrays (ospray_cpp/Data.h of OSPRay 3.x).
// SharedData (Data<true>) refers to the application's memory, CopiedData
// (Data<false>) copies it into an array the device owns; MovedData takes
// over a std::vector and frees it with the array.  Element types come from
// OSPTypeFor; dimensions are an integer (1D) or a 2- or 3-component vector
// of integers registered as OSP_VEC2UL / OSP_VEC3UL (or std::array).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

namespace detail {
template <typename DIM_T, typename = void>
struct Dims
{
  // a 2- or 3-component vector type with operator[]
  static void get(const DIM_T &d, uint64_t n[3])
  {
    constexpr OSPDataType t = OSPTypeFor<DIM_T>::value;
    static_assert(t == OSP_VEC2UL || t == OSP_VEC3UL || t == OSP_VEC2L || t == OSP_VEC3L || t == OSP_VEC2UI
            || t == OSP_VEC3UI || t == OSP_VEC2I || t == OSP_VEC3I,
        "Data dimensions: an integer or a 2/3-component integer vector");
    constexpr int c = (t == OSP_VEC2UL || t == OSP_VEC2L || t == OSP_VEC2UI || t == OSP_VEC2I) ? 2 : 3;
    for (int i = 0; i < 3; i++)
      n[i] = i < c ? (uint64_t)d[i] : 1;
  }
};
template <typename DIM_T>
struct Dims<DIM_T, typename std::enable_if<std::is_integral<DIM_T>::value>::type>
{
  static void get(const DIM_T &d, uint64_t n[3])
  {
    n[0] = (uint64_t)d;
    n[1] = n[2] = 1;
  }
};
template <typename I, std::size_t N>
struct Dims<std::array<I, N>, void>
{
  static void get(const std::array<I, N> &d, uint64_t n[3])
  {
    static_assert(N >= 1 && N <= 3, "Data dimensions: 1 to 3");
    for (std::size_t i = 0; i < 3; i++)
      n[i] = i < N ? (uint64_t)d[i] : 1;
  }
};
} // namespace detail

template <bool SHARED = false>
class Data : public ManagedObject<OSPData, OSP_DATA>
{
 public:
  template <typename DIM_T>
  Data(const void *init, OSPDataType type, const DIM_T &numItems)
  {
    uint64_t n[3];
    detail::Dims<DIM_T>::get(numItems, n);
    construct(init, type, n, nullptr);
  }
  template <typename DIM_T>
  Data(const void *init, OSPDataType type, const DIM_T &numItems, const DIM_T &byteStride)
  {
    uint64_t n[3], s[3];
    detail::Dims<DIM_T>::get(numItems, n);
    detail::Dims<DIM_T>::get(byteStride, s);
    const int64_t st[3] = {(int64_t)s[0], (int64_t)s[1], (int64_t)s[2]};
    construct(init, type, n, st);
  }
  template <typename T, typename DIM_T>
  Data(const T *init, const DIM_T &numItems) : Data((const void *)init, element<T>(), numItems)
  {
  }
  template <typename T, typename DIM_T>
  Data(const T *init, const DIM_T &numItems, const DIM_T &byteStride)
      : Data((const void *)init, element<T>(), numItems, byteStride)
  {
  }
  template <typename T, std::size_t N>
  Data(const std::array<T, N> &arr) : Data(arr.data(), N)
  {
  }
  template <typename T, typename ALLOC_T>
  Data(const std::vector<T, ALLOC_T> &arr) : Data(arr.data(), arr.size())
  {
  }
  // a single object or value as a 1-item array
  template <typename T>
  explicit Data(const T &obj) : Data(&obj, size_t(1))
  {
  }
  Data(OSPData existing = nullptr) : ManagedObject<OSPData, OSP_DATA>(existing) {}

 protected:
  template <typename T>
  static constexpr OSPDataType element()
  {
    static_assert(OSPTypeFor<T>::value != OSP_UNKNOWN, "Data: the element type has no OSPDataType");
    return OSPTypeFor<T>::value;
  }

 private:
  void construct(const void *init, OSPDataType type, const uint64_t n[3], const int64_t *stride)
  {
    const int64_t s0 = stride ? stride[0] : 0, s1 = stride ? stride[1] : 0, s2 = stride ? stride[2] : 0;
    OSPData shared = ospNewSharedData(init, type, n[0], s0, n[1], s1, n[2], s2, nullptr, nullptr);
    if (SHARED) {
      ospObject = shared;
      return;
    }
    ospObject = ospNewData(type, n[0], n[1], n[2]);
    ospCopyData(shared, ospObject, 0, 0, 0);
    ospRelease(shared);
  }
};

using SharedData = Data<true>;
using CopiedData = Data<false>;

// shares a std::vector's memory and frees the vector with the array
class MovedData : public SharedData
{
 public:
  template <typename T, typename ALLOC_T>
  MovedData(std::vector<T, ALLOC_T> &arr)
  {
    using V = std::vector<T, ALLOC_T>;
    V *owned = new V(std::move(arr));
    ospObject = ospNewSharedData(owned->data(), element<T>(), owned->size(), 0, 1, 0, 1, 0, &MovedData::free<V>,
        owned);
  }

 private:
  template <typename V>
  static void free(const void *userData, const void *)
  {
    delete (const V *)userData;
  }
};

} // namespace cpp

template <bool SHARED>
struct OSPTypeFor<cpp::Data<SHARED>>
{
  static constexpr OSPDataType value = OSP_DATA;
};
OSPTYPEFOR_SPECIALIZATION(cpp::MovedData, OSP_DATA);

} // namespace ospray
