// Internal C++ interface between libospray.so (the C API front end) and
// device modules such as libospray_module_hip.so.
//
// Plays the role of the reference's api::Device vtable
// (ospray/api/Device.h:23-146), ManagedObject parameter store
// (ospray/common/Managed.h:21-149) and module/device registry
// (ospray/common/OSPCommon.cpp:749-782, ospray/api/Device.cpp:33-45).
// Devices throw std::runtime_error; the C API front end converts the
// exception into an OSPError and the function's default return value
// (ospray/api/API.cpp:45-66).
#pragma once

#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "ospray/ospray.h"

namespace ospcore {

struct ManagedObject;

size_t sizeOfDataType(OSPDataType t);
bool isObjectType(OSPDataType t);
const char *dataTypeName(OSPDataType t);

struct Param
{
  OSPDataType type = OSP_UNKNOWN;
  alignas(16) unsigned char bytes[64];
  std::string str;
  ManagedObject *obj = nullptr;
  bool queried = false;
};

struct ManagedObject
{
  std::atomic<int64_t> refCount{1};
  OSPDataType managedType = OSP_OBJECT;
  std::string subtype;
  std::map<std::string, Param> params;
  uint64_t commitCounter = 0; // bumped on every commit

  ManagedObject() = default;
  ManagedObject(const ManagedObject &) = delete;
  ManagedObject &operator=(const ManagedObject &) = delete;
  virtual ~ManagedObject();

  void refInc() { refCount.fetch_add(1); }
  void refDec()
  {
    if (refCount.fetch_sub(1) == 1)
      delete this;
  }

  virtual void commit() {}
  virtual std::string toString() const { return "ospray::ManagedObject"; }

  void setParam(const char *name, OSPDataType type, const void *mem);
  void removeParam(const char *name);
  const Param *findParam(const char *name, OSPDataType type) const;
  Param *findParamAny(const char *name);

  // typed getters: the stored type must match exactly, as in the reference
  // (Managed.h getParam<T>); a parameter of another type is ignored
  float getFloat(const char *name, float def);
  int getInt(const char *name, int def);
  uint32_t getUInt(const char *name, uint32_t def);
  bool getBool(const char *name, bool def);
  void getVec3f(const char *name, float out[3], const float def[3]);
  void getVec4f(const char *name, float out[4], const float def[4]);
  bool getAffine3f(const char *name, float out[12]);
  std::string getString(const char *name, const std::string &def);
  ManagedObject *getObject(const char *name, OSPDataType type);
  std::vector<std::string> unusedParams() const;
};

struct Data : public ManagedObject
{
  OSPDataType elemType = OSP_UNKNOWN;
  uint64_t numItems[3] = {1, 1, 1};
  int64_t byteStride[3] = {0, 0, 0};
  const unsigned char *shared = nullptr;
  std::vector<unsigned char> owned;
  OSPDeleterCallback deleter = nullptr;
  const void *deleterUserPtr = nullptr;
  std::vector<ManagedObject *> objRefs; // refs held for object arrays

  Data() { managedType = OSP_DATA; }
  ~Data() override;
  uint64_t size() const { return numItems[0] * numItems[1] * numItems[2]; }
  const unsigned char *base() const { return shared ? shared : owned.data(); }
  bool compact() const;
  const unsigned char *item(uint64_t i) const;
  // element i (compact copy when strided)
  template <typename T>
  T get(uint64_t i) const
  {
    T v;
    std::memcpy(&v, item(i), sizeof(T));
    return v;
  }
  std::string toString() const override { return "ospray::Data"; }
};

struct Future;

class Device : public ManagedObject
{
 public:
  OSPError lastErrorCode = OSP_NO_ERROR;
  std::string lastErrorMsg;
  OSPErrorCallback errorFcn = nullptr;
  void *errorUserData = nullptr;
  OSPStatusCallback statusFcn = nullptr;
  void *statusUserData = nullptr;
  int logLevel = OSP_LOG_WARNING;
  bool committed = false;

  Device() { managedType = OSP_DEVICE; }
  ~Device() override = default;

  void commit() override;
  virtual void commitDevice() {}

  virtual OSPData newSharedData(const void *mem, OSPDataType, uint64_t n1, int64_t s1, uint64_t n2, int64_t s2,
      uint64_t n3, int64_t s3, OSPDeleterCallback, const void *userPtr);
  virtual OSPData newData(OSPDataType, uint64_t n1, uint64_t n2, uint64_t n3);
  virtual void copyData(Data *src, Data *dst, uint64_t i1, uint64_t i2, uint64_t i3);

  // new objects of the given handle kind (OSP_CAMERA, OSP_GEOMETRY, ...)
  virtual ManagedObject *newObject(OSPDataType kind, const char *type) = 0;
  virtual void commitObject(ManagedObject *obj);
  virtual OSPBounds getBounds(ManagedObject *obj) = 0;

  virtual ManagedObject *frameBufferCreate(int w, int h, OSPFrameBufferFormat, uint32_t channels) = 0;
  virtual const void *frameBufferMap(ManagedObject *fb, OSPFrameBufferChannel) = 0;
  virtual void frameBufferUnmap(const void *mapped, ManagedObject *fb) = 0;
  virtual float getVariance(ManagedObject *fb) = 0;
  virtual void resetAccumulation(ManagedObject *fb) = 0;

  virtual ManagedObject *renderFrame(ManagedObject *fb, ManagedObject *renderer, ManagedObject *camera, ManagedObject *world) = 0;
  virtual int isReady(ManagedObject *future, OSPSyncEvent) = 0;
  virtual void wait(ManagedObject *future, OSPSyncEvent) = 0;
  virtual void cancel(ManagedObject *future) = 0;
  virtual float getProgress(ManagedObject *future) = 0;
  virtual float getTaskDuration(ManagedObject *future) = 0;
  virtual OSPPickResult pick(ManagedObject *fb, ManagedObject *r, ManagedObject *c, ManagedObject *w, float x, float y) = 0;

  virtual int64_t getProperty(OSPDeviceProperty p);

  void postStatus(int level, const std::string &msg);
  void warnOnce(const std::string &msg);

 private:
  std::vector<std::string> warned_;
};

using DeviceFactory = Device *(*)();

} // namespace ospcore

extern "C" {
// Registers a device type; called from a module's ospray_module_init_<name>.
OSPRAY_INTERFACE void ospray_register_device_type(const char *name, ospcore::DeviceFactory factory);
// The device that C API calls dispatch to (nullptr before ospInit).
OSPRAY_INTERFACE ospcore::Device *ospray_current_device();
}
