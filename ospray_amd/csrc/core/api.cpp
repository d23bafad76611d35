// OSPRay C API front end (libospray.so).
//
// Every entry point forwards to the current ospcore::Device.  Exceptions
// thrown below are turned into an OSPError, reported through the device's
// error callback and last-error state, and the function returns its default
// value -- the contract of OSPRAY_CATCH_BEGIN/END (ospray/api/API.cpp:45-66).
// Module loading follows ospray/common/OSPCommon.cpp:749-782: dlopen
// "libospray_module_<name>.so", then call
// "ospray_module_init_<name>(major, minor, patch)".
#include <dlfcn.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>

#include "core.h"
#include "ospray/ospray_util.h"

using namespace ospcore;

namespace {

Device *g_current = nullptr;
std::map<std::string, DeviceFactory> &registry()
{
  static std::map<std::string, DeviceFactory> r;
  return r;
}
std::map<std::string, void *> &loadedModules()
{
  static std::map<std::string, void *> m;
  return m;
}
OSPError g_lastInitError = OSP_NO_ERROR;
std::string g_lastInitMsg;

void handleError(OSPError e, const std::string &msg)
{
  if (g_current) {
    g_current->lastErrorCode = e;
    g_current->lastErrorMsg = msg;
    if (g_current->errorFcn)
      g_current->errorFcn(g_current->errorUserData, e, msg.c_str());
    else
      std::fprintf(stderr, "OSPRAY ERROR: %s\n", msg.c_str());
  } else {
    g_lastInitError = e;
    g_lastInitMsg = msg;
    std::fprintf(stderr, "OSPRAY ERROR: %s\n", msg.c_str());
  }
}

Device &currentDevice()
{
  if (!g_current)
    throw std::runtime_error(
        "no valid OSPRay device while calling an API function, likely either before calling ospInit() / "
        "ospNewDevice() plus ospSetCurrentDevice(), or after calling ospShutdown(); pid: "
        + std::to_string(getpid()));
  return *g_current;
}

#define THROW_IF_NULL(obj, name)                                                                                  \
  if ((obj) == nullptr)                                                                                            \
  throw std::runtime_error(std::string("null ") + name + " provided to " + __func__)

#define CATCH_BEGIN try {
#define CATCH_END(a)                                                                                               \
  }                                                                                                                \
  catch (const std::bad_alloc &)                                                                                   \
  {                                                                                                                \
    handleError(OSP_OUT_OF_MEMORY, "OSPRay was unable to allocate memory");                                       \
    return a;                                                                                                      \
  }                                                                                                                \
  catch (const std::exception &e)                                                                                  \
  {                                                                                                                \
    handleError(OSP_UNKNOWN_ERROR, e.what());                                                                      \
    return a;                                                                                                      \
  }                                                                                                                \
  catch (...)                                                                                                      \
  {                                                                                                                \
    handleError(OSP_UNKNOWN_ERROR, "an unrecognized exception was caught");                                      \
    return a;                                                                                                      \
  }

std::string libraryDir()
{
  Dl_info info;
  if (dladdr((void *)&ospLoadModule, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t s = p.rfind('/');
    if (s != std::string::npos)
      return p.substr(0, s + 1);
  }
  return std::string();
}

OSPError loadLocalModule(const std::string &name)
{
  if (loadedModules().count(name))
    return OSP_NO_ERROR;
  const std::string file = "libospray_module_" + name + ".so";
  void *h = nullptr;
  const std::string dir = libraryDir();
  if (!dir.empty())
    h = dlopen((dir + file).c_str(), RTLD_NOW | RTLD_GLOBAL);
  if (!h)
    h = dlopen(file.c_str(), RTLD_NOW | RTLD_GLOBAL);
  if (!h)
    throw std::runtime_error("could not load module '" + name + "' (" + file + "): " + std::string(dlerror()));
  const std::string sym = "ospray_module_init_" + name;
  using InitFn = OSPError (*)(int16_t, int16_t, int16_t);
  auto fn = (InitFn)dlsym(h, sym.c_str());
  if (!fn)
    throw std::runtime_error("module '" + name + "' has no entry point " + sym);
  const OSPError e = fn(OSPRAY_VERSION_MAJOR, OSPRAY_VERSION_MINOR, OSPRAY_VERSION_PATCH);
  if (e != OSP_NO_ERROR)
    return e;
  loadedModules()[name] = h;
  return OSP_NO_ERROR;
}

Device *createDevice(const std::string &name)
{
  auto it = registry().find(name);
  if (it == registry().end()) {
    // a device type lives in the module of the same name
    if (loadLocalModule(name) == OSP_NO_ERROR)
      it = registry().find(name);
  }
  if (it == registry().end())
    throw std::runtime_error("Could not find device of type: " + name);
  return it->second();
}

std::string normalizeType(const char *t)
{
  std::string s(t);
  for (auto &c : s)
    if (c == '-' || c == ':')
      c = '_';
  return s;
}

ManagedObject *obj(OSPObject o) { return (ManagedObject *)o; }

} // namespace

extern "C" {

void ospray_register_device_type(const char *name, DeviceFactory factory) { registry()[name] = factory; }

Device *ospray_current_device() { return g_current; }

// ---- initialisation -----------------------------------------------------------

OSPError ospInit(int *argc, const char **argv)
{
  CATCH_BEGIN
  if (g_current)
    throw std::runtime_error("device already exists [ospInit() called twice?]");
  if (const char *mods = std::getenv("OSPRAY_LOAD_MODULES")) {
    std::stringstream ss(mods);
    std::string m;
    while (std::getline(ss, m, ','))
      if (!m.empty())
        loadLocalModule(m);
  }
  Device *dev = nullptr;
  if (argc && argv) {
    for (int i = 1; i < *argc;) {
      const std::string av(argv[i]);
      auto value = [&]() { return av.substr(av.find('=') + 1); };
      bool consumed = false;
      if (av.rfind("--osp:load-modules", 0) == 0) {
        std::stringstream ss(value());
        std::string m;
        while (std::getline(ss, m, ','))
          if (!m.empty())
            loadLocalModule(m);
        consumed = true;
      } else if (av.rfind("--osp:device=", 0) == 0) {
        dev = createDevice(value());
        consumed = true;
      } else if (av.rfind("--osp:log-level=", 0) == 0 && dev) {
        const std::string v = value();
        dev->setParam("logLevel", OSP_STRING, v.c_str());
        consumed = true;
      }
      if (consumed) {
        for (int j = i; j + 1 < *argc; j++)
          argv[j] = argv[j + 1];
        (*argc)--;
      } else {
        i++;
      }
    }
  }
  if (!dev) {
    // This library ships one device, "hip"; it is the default where the
    // reference defaults to "cpu" (ospray/api/API.cpp:134-145).
    const char *envDev = std::getenv("OSPRAY_DEVICE");
    dev = createDevice(envDev ? envDev : "hip");
  }
  g_current = dev;
  dev->commit();
  return OSP_NO_ERROR;
  CATCH_END(g_lastInitError == OSP_NO_ERROR ? OSP_INVALID_OPERATION : g_lastInitError)
}

int64_t ospDeviceGetProperty(OSPDevice d, OSPDeviceProperty p)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  return ((Device *)d)->getProperty(p);
  CATCH_END(0)
}

void ospShutdown()
{
  CATCH_BEGIN
  if (g_current)
    g_current->refDec();
  g_current = nullptr;
  CATCH_END()
}

OSPDevice ospNewDevice(const char *type)
{
  CATCH_BEGIN
  THROW_IF_NULL(type, "device type");
  return (OSPDevice)createDevice(type);
  CATCH_END(nullptr)
}

void ospSetCurrentDevice(OSPDevice d)
{
  CATCH_BEGIN
  Device *dev = (Device *)d;
  if (dev && !dev->committed)
    throw std::runtime_error("You must commit the device before using it!");
  if (dev)
    dev->refInc();
  if (g_current)
    g_current->refDec();
  g_current = dev;
  CATCH_END()
}

OSPDevice ospGetCurrentDevice()
{
  CATCH_BEGIN
  if (g_current)
    g_current->refInc();
  return (OSPDevice)g_current;
  CATCH_END(nullptr)
}

void ospDeviceSetParam(OSPDevice d, const char *id, OSPDataType type, const void *mem)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  THROW_IF_NULL(id, "parameter name");
  switch (type) {
  case OSP_STRING:
  case OSP_INT:
  case OSP_UINT:
  case OSP_BOOL:
    ((Device *)d)->setParam(id, type, mem);
    break;
  case OSP_VOID_PTR:
    ((Device *)d)->setParam(id, type, &mem);
    break;
  default:
    throw std::runtime_error("parameter type not handled for OSPDevice!");
  }
  CATCH_END()
}

void ospDeviceRemoveParam(OSPDevice d, const char *id)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  THROW_IF_NULL(id, "parameter name");
  ((Device *)d)->removeParam(id);
  CATCH_END()
}

void ospDeviceSetStatusCallback(OSPDevice d, OSPStatusCallback cb, void *user)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  ((Device *)d)->statusFcn = cb;
  ((Device *)d)->statusUserData = user;
  CATCH_END()
}

void ospDeviceSetErrorCallback(OSPDevice d, OSPErrorCallback cb, void *user)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  ((Device *)d)->errorFcn = cb;
  ((Device *)d)->errorUserData = user;
  CATCH_END()
}

OSPError ospDeviceGetLastErrorCode(OSPDevice d)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  return ((Device *)d)->lastErrorCode;
  CATCH_END(OSP_NO_ERROR)
}

const char *ospDeviceGetLastErrorMsg(OSPDevice d)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  return ((Device *)d)->lastErrorMsg.c_str();
  CATCH_END(nullptr)
}

void ospDeviceCommit(OSPDevice d)
{
  CATCH_BEGIN
  THROW_IF_NULL(d, "device");
  ((Device *)d)->commit();
  CATCH_END()
}

void ospDeviceRelease(OSPDevice d)
{
  CATCH_BEGIN
  if (!d)
    return;
  Device *dev = (Device *)d;
  if (dev == g_current && dev->refCount.load() == 1)
    g_current = nullptr;
  dev->refDec();
  CATCH_END()
}

void ospDeviceRetain(OSPDevice d)
{
  CATCH_BEGIN
  if (d)
    ((Device *)d)->refInc();
  CATCH_END()
}

OSPError ospLoadModule(const char *name)
{
  CATCH_BEGIN
  THROW_IF_NULL(name, "module name");
  return loadLocalModule(name);
  CATCH_END(OSP_UNKNOWN_ERROR)
}

// ---- data -------------------------------------------------------------------------

OSPData ospNewSharedData(const void *mem, OSPDataType type, uint64_t n1, int64_t s1, uint64_t n2, int64_t s2,
    uint64_t n3, int64_t s3, OSPDeleterCallback del, const void *user)
{
  CATCH_BEGIN
  return currentDevice().newSharedData(mem, type, n1, s1, n2, s2, n3, s3, del, user);
  CATCH_END(nullptr)
}

OSPData ospNewData(OSPDataType type, uint64_t n1, uint64_t n2, uint64_t n3)
{
  CATCH_BEGIN
  return currentDevice().newData(type, n1, n2, n3);
  CATCH_END(nullptr)
}

void ospCopyData(const OSPData src, OSPData dst, uint64_t i1, uint64_t i2, uint64_t i3)
{
  CATCH_BEGIN
  THROW_IF_NULL(src, "source data");
  THROW_IF_NULL(dst, "destination data");
  currentDevice().copyData((Data *)src, (Data *)dst, i1, i2, i3);
  CATCH_END()
}

// ---- objects -------------------------------------------------------------------------

#define NEW_OBJECT(fn, handle, kind)                                                                               \
  handle fn(const char *type)                                                                                      \
  {                                                                                                                \
    CATCH_BEGIN                                                                                                    \
    THROW_IF_NULL(type, "type string");                                                                            \
    return (handle)currentDevice().newObject(kind, normalizeType(type).c_str());                                  \
    CATCH_END(nullptr)                                                                                             \
  }

NEW_OBJECT(ospNewLight, OSPLight, OSP_LIGHT)
NEW_OBJECT(ospNewCamera, OSPCamera, OSP_CAMERA)
NEW_OBJECT(ospNewGeometry, OSPGeometry, OSP_GEOMETRY)
NEW_OBJECT(ospNewVolume, OSPVolume, OSP_VOLUME)
NEW_OBJECT(ospNewMaterial, OSPMaterial, OSP_MATERIAL)
NEW_OBJECT(ospNewTransferFunction, OSPTransferFunction, OSP_TRANSFER_FUNCTION)
NEW_OBJECT(ospNewTexture, OSPTexture, OSP_TEXTURE)
NEW_OBJECT(ospNewRenderer, OSPRenderer, OSP_RENDERER)
NEW_OBJECT(ospNewImageOperation, OSPImageOperation, OSP_IMAGE_OPERATION)

OSPGeometricModel ospNewGeometricModel(OSPGeometry geom)
{
  CATCH_BEGIN
  ManagedObject *m = currentDevice().newObject(OSP_GEOMETRIC_MODEL, "");
  if (geom)
    m->setParam("geometry", OSP_GEOMETRY, &geom);
  return (OSPGeometricModel)m;
  CATCH_END(nullptr)
}

OSPVolumetricModel ospNewVolumetricModel(OSPVolume vol)
{
  CATCH_BEGIN
  ManagedObject *m = currentDevice().newObject(OSP_VOLUMETRIC_MODEL, "");
  if (vol)
    m->setParam("volume", OSP_VOLUME, &vol);
  return (OSPVolumetricModel)m;
  CATCH_END(nullptr)
}

OSPGroup ospNewGroup()
{
  CATCH_BEGIN
  return (OSPGroup)currentDevice().newObject(OSP_GROUP, "");
  CATCH_END(nullptr)
}

OSPInstance ospNewInstance(OSPGroup group)
{
  CATCH_BEGIN
  ManagedObject *i = currentDevice().newObject(OSP_INSTANCE, "");
  if (group)
    i->setParam("group", OSP_GROUP, &group);
  return (OSPInstance)i;
  CATCH_END(nullptr)
}

OSPWorld ospNewWorld()
{
  CATCH_BEGIN
  return (OSPWorld)currentDevice().newObject(OSP_WORLD, "");
  CATCH_END(nullptr)
}

OSPBounds ospGetBounds(OSPObject o)
{
  OSPBounds empty;
  for (int i = 0; i < 3; i++) {
    empty.lower[i] = INFINITY;
    empty.upper[i] = -INFINITY;
  }
  CATCH_BEGIN
  THROW_IF_NULL(o, "object");
  return currentDevice().getBounds(obj(o));
  CATCH_END(empty)
}

void ospSetParam(OSPObject o, const char *id, OSPDataType type, const void *mem)
{
  CATCH_BEGIN
  THROW_IF_NULL(o, "object");
  THROW_IF_NULL(id, "parameter name");
  THROW_IF_NULL(mem, "parameter value");
  obj(o)->setParam(id, type, mem);
  CATCH_END()
}

void ospRemoveParam(OSPObject o, const char *id)
{
  CATCH_BEGIN
  THROW_IF_NULL(o, "object");
  THROW_IF_NULL(id, "parameter name");
  obj(o)->removeParam(id);
  CATCH_END()
}

void ospCommit(OSPObject o)
{
  CATCH_BEGIN
  THROW_IF_NULL(o, "object");
  currentDevice().commitObject(obj(o));
  CATCH_END()
}

void ospRelease(OSPObject o)
{
  CATCH_BEGIN
  if (o)
    obj(o)->refDec();
  CATCH_END()
}

void ospRetain(OSPObject o)
{
  CATCH_BEGIN
  if (o)
    obj(o)->refInc();
  CATCH_END()
}

// ---- framebuffer -------------------------------------------------------------------------

OSPFrameBuffer ospNewFrameBuffer(int w, int h, OSPFrameBufferFormat format, uint32_t channels)
{
  CATCH_BEGIN
  uint32_t ch = channels;
  if ((channels & OSP_FB_ACCUM) == 0)
    ch &= ~(uint32_t)OSP_FB_VARIANCE;
  return (OSPFrameBuffer)currentDevice().frameBufferCreate(w, h, format, ch);
  CATCH_END(nullptr)
}

const void *ospMapFrameBuffer(OSPFrameBuffer fb, OSPFrameBufferChannel ch)
{
  CATCH_BEGIN
  THROW_IF_NULL(fb, "framebuffer");
  return currentDevice().frameBufferMap(obj(fb), ch);
  CATCH_END(nullptr)
}

void ospUnmapFrameBuffer(const void *mapped, OSPFrameBuffer fb)
{
  CATCH_BEGIN
  THROW_IF_NULL(fb, "framebuffer");
  THROW_IF_NULL(mapped, "pointer");
  currentDevice().frameBufferUnmap(mapped, obj(fb));
  CATCH_END()
}

float ospGetVariance(OSPFrameBuffer fb)
{
  CATCH_BEGIN
  THROW_IF_NULL(fb, "framebuffer");
  return currentDevice().getVariance(obj(fb));
  CATCH_END(INFINITY)
}

void ospResetAccumulation(OSPFrameBuffer fb)
{
  CATCH_BEGIN
  THROW_IF_NULL(fb, "framebuffer");
  currentDevice().resetAccumulation(obj(fb));
  CATCH_END()
}

// ---- rendering -----------------------------------------------------------------------------

OSPFuture ospRenderFrame(OSPFrameBuffer fb, OSPRenderer r, OSPCamera c, OSPWorld w)
{
  CATCH_BEGIN
  THROW_IF_NULL(fb, "framebuffer");
  THROW_IF_NULL(r, "renderer");
  THROW_IF_NULL(c, "camera");
  THROW_IF_NULL(w, "world");
  return (OSPFuture)currentDevice().renderFrame(obj(fb), obj(r), obj(c), obj(w));
  CATCH_END(nullptr)
}

int ospIsReady(OSPFuture f, OSPSyncEvent e)
{
  CATCH_BEGIN
  THROW_IF_NULL(f, "future");
  return currentDevice().isReady(obj(f), e);
  CATCH_END(1)
}

void ospWait(OSPFuture f, OSPSyncEvent e)
{
  CATCH_BEGIN
  THROW_IF_NULL(f, "future");
  if (e == OSP_NONE_FINISHED)
    return;
  currentDevice().wait(obj(f), e);
  CATCH_END()
}

void ospCancel(OSPFuture f)
{
  CATCH_BEGIN
  THROW_IF_NULL(f, "future");
  currentDevice().cancel(obj(f));
  CATCH_END()
}

float ospGetProgress(OSPFuture f)
{
  CATCH_BEGIN
  THROW_IF_NULL(f, "future");
  return currentDevice().getProgress(obj(f));
  CATCH_END(1.f)
}

float ospGetTaskDuration(OSPFuture f)
{
  CATCH_BEGIN
  THROW_IF_NULL(f, "future");
  return currentDevice().getTaskDuration(obj(f));
  CATCH_END(0.f)
}

void ospPick(OSPPickResult *result, OSPFrameBuffer fb, OSPRenderer r, OSPCamera c, OSPWorld w, float x, float y)
{
  CATCH_BEGIN
  THROW_IF_NULL(result, "pick result");
  THROW_IF_NULL(fb, "framebuffer");
  THROW_IF_NULL(r, "renderer");
  THROW_IF_NULL(c, "camera");
  THROW_IF_NULL(w, "world");
  *result = currentDevice().pick(obj(fb), obj(r), obj(c), obj(w), x, y);
  CATCH_END()
}

} // extern "C"
