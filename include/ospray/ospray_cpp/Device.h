// OSPRay C++ wrapper: Device (ospray_cpp/Device.h of OSPRay 3.x).  A
// device is not an OSPObject: it is retained and released with
// ospDeviceRetain / ospDeviceRelease.
#pragma once

#include <string>
#include <utility>

#include "ospray/ospray.h"

namespace ospray {
namespace cpp {

class Device
{
 public:
  Device(const std::string &type = "cpu") : ospHandle(ospNewDevice(type.c_str())) {}
  Device(const Device &copy) : ospHandle(copy.ospHandle)
  {
    if (ospHandle)
      ospDeviceRetain(ospHandle);
  }
  Device(Device &&move) : ospHandle(move.ospHandle) { move.ospHandle = nullptr; }
  Device(OSPDevice existing) : ospHandle(existing) {}
  ~Device()
  {
    if (ospHandle)
      ospDeviceRelease(ospHandle);
  }
  Device &operator=(const Device &copy)
  {
    if (copy.ospHandle)
      ospDeviceRetain(copy.ospHandle);
    if (ospHandle)
      ospDeviceRelease(ospHandle);
    ospHandle = copy.ospHandle;
    return *this;
  }
  Device &operator=(Device &&move)
  {
    if (this != &move) {
      if (ospHandle)
        ospDeviceRelease(ospHandle);
      ospHandle = move.ospHandle;
      move.ospHandle = nullptr;
    }
    return *this;
  }

  // a callable that converts to OSPErrorCallback / OSPStatusCallback (a
  // function or a lambda without captures); no user data
  template <typename FCN_T>
  void setErrorCallback(FCN_T &&fcn) const
  {
    ospDeviceSetErrorCallback(ospHandle, std::forward<FCN_T>(fcn), nullptr);
  }
  template <typename FCN_T>
  void setStatusCallback(FCN_T &&fcn) const
  {
    ospDeviceSetStatusCallback(ospHandle, std::forward<FCN_T>(fcn), nullptr);
  }

  void setParam(const std::string &name, const int v) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), OSP_INT, &v);
  }
  void setParam(const std::string &name, const unsigned int v) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), OSP_UINT, &v);
  }
  void setParam(const std::string &name, const bool v) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), OSP_BOOL, &v);
  }
  void setParam(const std::string &name, const void *v) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), OSP_VOID_PTR, &v);
  }
  void setParam(const std::string &name, const char *v) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), OSP_STRING, v);
  }
  void setParam(const std::string &name, const std::string &v) const { setParam(name, v.c_str()); }
  void setParam(const std::string &name, OSPDataType t, const void *mem) const
  {
    ospDeviceSetParam(ospHandle, name.c_str(), t, mem);
  }
  void removeParam(const std::string &name) const { ospDeviceRemoveParam(ospHandle, name.c_str()); }
  void commit() const { ospDeviceCommit(ospHandle); }
  void setCurrent() const { ospSetCurrentDevice(ospHandle); }
  OSPDevice handle() const { return ospHandle; }
  static Device current() { return Device(ospGetCurrentDevice()); } // ospGetCurrentDevice retains

 private:
  OSPDevice ospHandle;
};

} // namespace cpp
} // namespace ospray
