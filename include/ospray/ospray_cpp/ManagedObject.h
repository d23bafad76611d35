// OSPRay C++ wrapper: ManagedObject, the owning handle every object class
// derives from (ospray_cpp/ManagedObject.h of OSPRay 3.x).  Copies retain
// the object, destruction releases it; setParam maps a C++ value to its
// OSPDataType through OSPTypeFor.
#pragma once

#include <string>
#include <type_traits>

#include "ospray/ospray.h"
#include "Traits.h"

namespace ospray {
namespace cpp {

template <typename HANDLE_T = OSPObject, OSPDataType TYPE = OSP_OBJECT>
class ManagedObject
{
 public:
  static constexpr OSPDataType type = TYPE;

  ManagedObject(HANDLE_T object = nullptr) : ospObject(object)
  {
    static_assert(std::is_same<HANDLE_T, OSPObject>::value
            || std::is_base_of<osp::ManagedObject, typename std::remove_pointer<HANDLE_T>::type>::value,
        "ManagedObject wraps OSPObject or one of the OSP* handle types");
  }
  ~ManagedObject() { ospRelease(ospObject); }

  ManagedObject(const ManagedObject &copy) : ospObject(copy.ospObject)
  {
    if (ospObject)
      ospRetain(ospObject);
  }
  ManagedObject(ManagedObject &&move) : ospObject(move.ospObject) { move.ospObject = nullptr; }
  ManagedObject &operator=(const ManagedObject &copy)
  {
    if (copy.ospObject)
      ospRetain(copy.ospObject);
    ospRelease(ospObject);
    ospObject = copy.ospObject;
    return *this;
  }
  ManagedObject &operator=(ManagedObject &&move)
  {
    if (this != &move) {
      ospRelease(ospObject);
      ospObject = move.ospObject;
      move.ospObject = nullptr;
    }
    return *this;
  }

  bool operator==(const ManagedObject &other) const { return ospObject == other.ospObject; }
  bool operator!=(const ManagedObject &other) const { return ospObject != other.ospObject; }

  template <typename T>
  void setParam(const std::string &name, const T &v) const
  {
    static_assert(OSPTypeFor<T>::value != OSP_UNKNOWN,
        "setParam: the value's type has no OSPDataType (register it with OSPTYPEFOR_SPECIALIZATION)");
    ospSetParam(ospObject, name.c_str(), OSPTypeFor<T>::value, &v);
  }
  void setParam(const std::string &name, const char *v) const { ospSetParam(ospObject, name.c_str(), OSP_STRING, v); }
  void setParam(const std::string &name, const std::string &v) const
  {
    ospSetParam(ospObject, name.c_str(), OSP_STRING, v.c_str());
  }
  void setParam(const std::string &name, OSPDataType t, const void *mem) const
  {
    ospSetParam(ospObject, name.c_str(), t, mem);
  }
  void removeParam(const std::string &name) const { ospRemoveParam(ospObject, name.c_str()); }

  // bounds as any box type laid out as lower[3], upper[3]
  template <typename BOX3F_T>
  BOX3F_T getBounds() const
  {
    static_assert(sizeof(BOX3F_T) == sizeof(OSPBounds), "getBounds: the box type must be 6 floats");
    const OSPBounds b = ospGetBounds(ospObject);
    BOX3F_T out;
    static_assert(std::is_trivially_copyable<BOX3F_T>::value, "getBounds: the box type must be trivially copyable");
    __builtin_memcpy(&out, &b, sizeof(b));
    return out;
  }

  void commit() const { ospCommit(ospObject); }
  HANDLE_T handle() const { return ospObject; }
  operator bool() const { return ospObject != nullptr; }

 protected:
  HANDLE_T ospObject;
};

} // namespace cpp

// a wrapper object passed to setParam is its handle
template <typename HANDLE_T, OSPDataType TYPE>
struct OSPTypeFor<cpp::ManagedObject<HANDLE_T, TYPE>>
{
  static constexpr OSPDataType value = TYPE;
};

} // namespace ospray
