// Parameter store, data arrays and device defaults shared by all devices.
#include <algorithm>
#include <cstdio>

#include "core.h"

namespace ospcore {

size_t sizeOfDataType(OSPDataType t)
{
  if (isObjectType(t))
    return sizeof(void *);
  switch (t) {
  case OSP_DEVICE:
  case OSP_VOID_PTR:
  case OSP_STRING:
    return sizeof(void *);
  case OSP_BOOL:
    return 1;
  case OSP_CHAR:
  case OSP_UCHAR:
    return 1;
  case OSP_VEC2C:
  case OSP_VEC2UC:
    return 2;
  case OSP_VEC3C:
  case OSP_VEC3UC:
    return 3;
  case OSP_VEC4C:
  case OSP_VEC4UC:
    return 4;
  case OSP_SHORT:
  case OSP_USHORT:
  case OSP_HALF:
    return 2;
  case OSP_VEC2S:
  case OSP_VEC2US:
  case OSP_VEC2H:
    return 4;
  case OSP_VEC3S:
  case OSP_VEC3US:
  case OSP_VEC3H:
    return 6;
  case OSP_VEC4S:
  case OSP_VEC4US:
  case OSP_VEC4H:
    return 8;
  case OSP_INT:
  case OSP_UINT:
  case OSP_FLOAT:
    return 4;
  case OSP_VEC2I:
  case OSP_VEC2UI:
  case OSP_VEC2F:
  case OSP_BOX1I:
  case OSP_BOX1F:
    return 8;
  case OSP_VEC3I:
  case OSP_VEC3UI:
  case OSP_VEC3F:
    return 12;
  case OSP_VEC4I:
  case OSP_VEC4UI:
  case OSP_VEC4F:
  case OSP_BOX2I:
  case OSP_BOX2F:
  case OSP_LINEAR2F:
  case OSP_QUATF:
    return 16;
  case OSP_LONG:
  case OSP_ULONG:
  case OSP_DOUBLE:
    return 8;
  case OSP_VEC2L:
  case OSP_VEC2UL:
  case OSP_VEC2D:
    return 16;
  case OSP_VEC3L:
  case OSP_VEC3UL:
  case OSP_VEC3D:
    return 24;
  case OSP_VEC4L:
  case OSP_VEC4UL:
  case OSP_VEC4D:
    return 32;
  case OSP_BOX3I:
  case OSP_BOX3F:
    return 24;
  case OSP_BOX4I:
  case OSP_BOX4F:
    return 32;
  case OSP_LINEAR3F:
    return 36;
  case OSP_AFFINE2F:
    return 24;
  case OSP_AFFINE3F:
    return 48;
  default:
    return 0;
  }
}

bool isObjectType(OSPDataType t) { return (uint32_t)t >= (uint32_t)OSP_OBJECT && (uint32_t)t <= (uint32_t)OSP_WORLD; }

const char *dataTypeName(OSPDataType t)
{
  switch (t) {
  case OSP_FLOAT: return "float";
  case OSP_INT: return "int";
  case OSP_UINT: return "uint";
  case OSP_BOOL: return "bool";
  case OSP_VEC3F: return "vec3f";
  case OSP_VEC4F: return "vec4f";
  case OSP_STRING: return "string";
  case OSP_DATA: return "data";
  default: return "value";
  }
}

ManagedObject::~ManagedObject()
{
  for (auto &kv : params)
    if (kv.second.obj)
      kv.second.obj->refDec();
}

void ManagedObject::setParam(const char *name, OSPDataType type, const void *mem)
{
  if (!name)
    throw std::runtime_error("parameter name is null");
  Param p;
  p.type = type;
  std::memset(p.bytes, 0, sizeof(p.bytes));
  if (type == OSP_STRING) {
    p.str = mem ? std::string((const char *)mem) : std::string();
  } else if (isObjectType(type)) {
    ManagedObject *o = mem ? *(ManagedObject *const *)mem : nullptr;
    if (o)
      o->refInc();
    p.obj = o;
  } else {
    const size_t sz = sizeOfDataType(type);
    if (sz == 0 || sz > sizeof(p.bytes))
      throw std::runtime_error(std::string("unsupported parameter type for '") + name + "'");
    if (type == OSP_BOOL) {
      // OSP_BOOL parameters are passed as int-sized values
      const int v = *(const unsigned char *)mem != 0;
      std::memcpy(p.bytes, &v, sizeof(int));
    } else {
      std::memcpy(p.bytes, mem, sz);
    }
  }
  auto it = params.find(name);
  if (it != params.end() && it->second.obj)
    it->second.obj->refDec();
  params[name] = p;
}

void ManagedObject::removeParam(const char *name)
{
  auto it = params.find(name);
  if (it == params.end())
    return;
  if (it->second.obj)
    it->second.obj->refDec();
  params.erase(it);
}

const Param *ManagedObject::findParam(const char *name, OSPDataType type) const
{
  auto it = params.find(name);
  if (it == params.end() || it->second.type != type)
    return nullptr;
  const_cast<Param &>(it->second).queried = true;
  return &it->second;
}

Param *ManagedObject::findParamAny(const char *name)
{
  auto it = params.find(name);
  if (it == params.end())
    return nullptr;
  it->second.queried = true;
  return &it->second;
}

float ManagedObject::getFloat(const char *name, float def)
{
  const Param *p = findParam(name, OSP_FLOAT);
  float v = def;
  if (p)
    std::memcpy(&v, p->bytes, 4);
  return v;
}

int ManagedObject::getInt(const char *name, int def)
{
  const Param *p = findParam(name, OSP_INT);
  int v = def;
  if (p)
    std::memcpy(&v, p->bytes, 4);
  return v;
}

uint32_t ManagedObject::getUInt(const char *name, uint32_t def)
{
  const Param *p = findParam(name, OSP_UINT);
  uint32_t v = def;
  if (p)
    std::memcpy(&v, p->bytes, 4);
  return v;
}

bool ManagedObject::getBool(const char *name, bool def)
{
  const Param *p = findParam(name, OSP_BOOL);
  if (!p)
    return def;
  int v;
  std::memcpy(&v, p->bytes, 4);
  return v != 0;
}

void ManagedObject::getVec3f(const char *name, float out[3], const float def[3])
{
  const Param *p = findParam(name, OSP_VEC3F);
  std::memcpy(out, p ? (const void *)p->bytes : (const void *)def, 12);
}

void ManagedObject::getVec4f(const char *name, float out[4], const float def[4])
{
  const Param *p = findParam(name, OSP_VEC4F);
  std::memcpy(out, p ? (const void *)p->bytes : (const void *)def, 16);
}

bool ManagedObject::getAffine3f(const char *name, float out[12])
{
  const Param *p = findParam(name, OSP_AFFINE3F);
  if (!p)
    return false;
  std::memcpy(out, p->bytes, 48);
  return true;
}

std::string ManagedObject::getString(const char *name, const std::string &def)
{
  const Param *p = findParam(name, OSP_STRING);
  return p ? p->str : def;
}

ManagedObject *ManagedObject::getObject(const char *name, OSPDataType type)
{
  // object handles may be set with their own type or as OSP_OBJECT
  // (ospSetObject); the object's own kind decides
  auto it = params.find(name);
  if (it == params.end() || !isObjectType(it->second.type) || !it->second.obj)
    return nullptr;
  ManagedObject *o = it->second.obj;
  if (type != OSP_OBJECT && o->managedType != type)
    return nullptr;
  it->second.queried = true;
  return o;
}

std::vector<std::string> ManagedObject::unusedParams() const
{
  std::vector<std::string> r;
  for (auto &kv : params)
    if (!kv.second.queried)
      r.push_back(kv.first);
  return r;
}

// ---- Data -------------------------------------------------------------------

Data::~Data()
{
  for (auto *o : objRefs)
    if (o)
      o->refDec();
  if (deleter && shared)
    deleter(deleterUserPtr, shared);
}

bool Data::compact() const
{
  const size_t es = sizeOfDataType(elemType);
  return byteStride[0] == (int64_t)es && byteStride[1] == (int64_t)(es * numItems[0])
      && byteStride[2] == (int64_t)(es * numItems[0] * numItems[1]);
}

const unsigned char *Data::item(uint64_t i) const
{
  const uint64_t i0 = i % numItems[0];
  const uint64_t r = i / numItems[0];
  const uint64_t i1 = r % numItems[1];
  const uint64_t i2 = r / numItems[1];
  return base() + (int64_t)i0 * byteStride[0] + (int64_t)i1 * byteStride[1] + (int64_t)i2 * byteStride[2];
}

// ---- Device -------------------------------------------------------------------

void Device::commit()
{
  // device parameters (ospray/api/Device.cpp:72-135)
  const std::string lvl = getString("logLevel", "");
  if (lvl == "debug")
    logLevel = OSP_LOG_DEBUG;
  else if (lvl == "info")
    logLevel = OSP_LOG_INFO;
  else if (lvl == "warning")
    logLevel = OSP_LOG_WARNING;
  else if (lvl == "error")
    logLevel = OSP_LOG_ERROR;
  else if (lvl == "none")
    logLevel = OSP_LOG_NONE;
  logLevel = getInt("logLevel", logLevel);
  commitDevice();
  committed = true;
}

void Device::postStatus(int level, const std::string &msg)
{
  if (level < logLevel)
    return;
  if (statusFcn)
    statusFcn(statusUserData, (msg + "\n").c_str());
  else
    std::fprintf(stderr, "%s\n", msg.c_str());
}

void Device::warnOnce(const std::string &msg)
{
  if (std::find(warned_.begin(), warned_.end(), msg) != warned_.end())
    return;
  warned_.push_back(msg);
  postStatus(OSP_LOG_WARNING, msg);
}

int64_t Device::getProperty(OSPDeviceProperty p)
{
  switch (p) {
  case OSP_DEVICE_VERSION:
    return 10000 * OSPRAY_VERSION_MAJOR + 100 * OSPRAY_VERSION_MINOR + OSPRAY_VERSION_PATCH;
  case OSP_DEVICE_VERSION_MAJOR:
    return OSPRAY_VERSION_MAJOR;
  case OSP_DEVICE_VERSION_MINOR:
    return OSPRAY_VERSION_MINOR;
  case OSP_DEVICE_VERSION_PATCH:
    return OSPRAY_VERSION_PATCH;
  case OSP_DEVICE_SO_VERSION:
    return 3;
  default:
    throw std::runtime_error("unknown device property");
  }
}

OSPData Device::newSharedData(const void *mem, OSPDataType type, uint64_t n1, int64_t s1, uint64_t n2, int64_t s2,
    uint64_t n3, int64_t s3, OSPDeleterCallback del, const void *userPtr)
{
  if (!mem)
    throw std::runtime_error("ospNewSharedData: shared memory is null");
  if (n1 == 0 || n2 == 0 || n3 == 0)
    throw std::runtime_error("ospNewSharedData: zero items");
  const size_t es = sizeOfDataType(type);
  if (es == 0)
    throw std::runtime_error("ospNewSharedData: unsupported element type");
  auto *d = new Data;
  d->elemType = type;
  d->numItems[0] = n1;
  d->numItems[1] = n2;
  d->numItems[2] = n3;
  d->byteStride[0] = s1 ? s1 : (int64_t)es;
  d->byteStride[1] = s2 ? s2 : d->byteStride[0] * (int64_t)n1;
  d->byteStride[2] = s3 ? s3 : d->byteStride[1] * (int64_t)n2;
  d->shared = (const unsigned char *)mem;
  d->deleter = del;
  d->deleterUserPtr = userPtr;
  if (isObjectType(type)) {
    // shared object arrays hold references (Data.cpp semantics)
    for (uint64_t i = 0; i < d->size(); i++) {
      ManagedObject *o = d->get<ManagedObject *>(i);
      if (o)
        o->refInc();
      d->objRefs.push_back(o);
    }
  }
  return (OSPData)d;
}

OSPData Device::newData(OSPDataType type, uint64_t n1, uint64_t n2, uint64_t n3)
{
  if (n1 == 0 || n2 == 0 || n3 == 0)
    throw std::runtime_error("ospNewData: zero items");
  const size_t es = sizeOfDataType(type);
  if (es == 0)
    throw std::runtime_error("ospNewData: unsupported element type");
  auto *d = new Data;
  d->elemType = type;
  d->numItems[0] = n1;
  d->numItems[1] = n2;
  d->numItems[2] = n3;
  d->byteStride[0] = (int64_t)es;
  d->byteStride[1] = (int64_t)(es * n1);
  d->byteStride[2] = (int64_t)(es * n1 * n2);
  d->owned.assign(es * n1 * n2 * n3, 0);
  if (isObjectType(type))
    d->objRefs.assign(d->size(), nullptr);
  return (OSPData)d;
}

void Device::copyData(Data *src, Data *dst, uint64_t i1, uint64_t i2, uint64_t i3)
{
  if (!src || !dst)
    throw std::runtime_error("ospCopyData: null array");
  if (dst->shared)
    throw std::runtime_error("ospCopyData: destination must be an array created with ospNewData");
  if (src->elemType != dst->elemType)
    throw std::runtime_error("ospCopyData: element types differ");
  if (i1 + src->numItems[0] > dst->numItems[0] || i2 + src->numItems[1] > dst->numItems[1]
      || i3 + src->numItems[2] > dst->numItems[2])
    throw std::runtime_error("ospCopyData: source does not fit destination region");
  const size_t es = sizeOfDataType(src->elemType);
  const bool objs = isObjectType(src->elemType);
  for (uint64_t z = 0; z < src->numItems[2]; z++)
    for (uint64_t y = 0; y < src->numItems[1]; y++)
      for (uint64_t x = 0; x < src->numItems[0]; x++) {
        const uint64_t si = x + src->numItems[0] * (y + src->numItems[1] * z);
        const uint64_t di = (x + i1) + dst->numItems[0] * ((y + i2) + dst->numItems[1] * (z + i3));
        unsigned char *dp = dst->owned.data() + di * es;
        if (objs) {
          ManagedObject *o = src->get<ManagedObject *>(si);
          if (o)
            o->refInc();
          if (dst->objRefs[di])
            dst->objRefs[di]->refDec();
          dst->objRefs[di] = o;
          std::memcpy(dp, &o, sizeof(o));
        } else {
          std::memcpy(dp, src->item(si), es);
        }
      }
}

void Device::commitObject(ManagedObject *obj)
{
  obj->commit();
  obj->commitCounter++;
}

} // namespace ospcore
