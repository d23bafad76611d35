// ospray_util.h helpers, written only in terms of the core C API
// (same behaviour as ospray/api/ospray_util_impl.cpp).
#include "ospray/ospray_util.h"

extern "C" {

OSPData ospNewSharedData1D(const void *p, OSPDataType t, uint64_t n)
{
  return ospNewSharedData(p, t, n, 0, 1, 0, 1, 0, NULL, NULL);
}

OSPData ospNewAssignedData1D(const void *p, OSPDataType t, uint64_t n, OSPDeleterCallback del, const void *user)
{
  return ospNewSharedData(p, t, n, 0, 1, 0, 1, 0, del, user);
}

OSPData ospNewSharedData1DStride(const void *p, OSPDataType t, uint64_t n, int64_t stride)
{
  return ospNewSharedData(p, t, n, stride, 1, 0, 1, 0, NULL, NULL);
}

OSPData ospNewSharedData2D(const void *p, OSPDataType t, uint64_t n1, uint64_t n2)
{
  return ospNewSharedData(p, t, n1, 0, n2, 0, 1, 0, NULL, NULL);
}

OSPData ospNewSharedData2DStride(const void *p, OSPDataType t, uint64_t n1, int64_t s1, uint64_t n2, int64_t s2)
{
  return ospNewSharedData(p, t, n1, s1, n2, s2, 1, 0, NULL, NULL);
}

OSPData ospNewSharedData3D(const void *p, OSPDataType t, uint64_t n1, uint64_t n2, uint64_t n3)
{
  return ospNewSharedData(p, t, n1, 0, n2, 0, n3, 0, NULL, NULL);
}

OSPData ospNewData1D(OSPDataType t, uint64_t n) { return ospNewData(t, n, 1, 1); }

OSPData ospNewData2D(OSPDataType t, uint64_t n1, uint64_t n2) { return ospNewData(t, n1, n2, 1); }

void ospCopyData1D(const OSPData src, OSPData dst, uint64_t i) { ospCopyData(src, dst, i, 0, 0); }

void ospCopyData2D(const OSPData src, OSPData dst, uint64_t i1, uint64_t i2) { ospCopyData(src, dst, i1, i2, 0); }

void ospSetString(OSPObject o, const char *n, const char *s) { ospSetParam(o, n, OSP_STRING, s); }

void ospSetObject(OSPObject o, const char *n, OSPObject obj) { ospSetParam(o, n, OSP_OBJECT, &obj); }

void ospSetBool(OSPObject o, const char *n, int x) { ospSetParam(o, n, OSP_BOOL, &x); }

void ospSetFloat(OSPObject o, const char *n, float x) { ospSetParam(o, n, OSP_FLOAT, &x); }

void ospSetInt(OSPObject o, const char *n, int x) { ospSetParam(o, n, OSP_INT, &x); }

void ospSetUInt(OSPObject o, const char *n, unsigned int x) { ospSetParam(o, n, OSP_UINT, &x); }

void ospSetBox1f(OSPObject o, const char *n, float x, float y)
{
  const float v[] = {x, y};
  ospSetParam(o, n, OSP_BOX1F, v);
}

void ospSetLinear2f(OSPObject o, const char *n, float x, float y, float z, float w)
{
  const float v[] = {x, y, z, w};
  ospSetParam(o, n, OSP_LINEAR2F, v);
}

void ospSetVec2f(OSPObject o, const char *n, float x, float y)
{
  const float v[] = {x, y};
  ospSetParam(o, n, OSP_VEC2F, v);
}

void ospSetVec3f(OSPObject o, const char *n, float x, float y, float z)
{
  const float v[] = {x, y, z};
  ospSetParam(o, n, OSP_VEC3F, v);
}

void ospSetVec4f(OSPObject o, const char *n, float x, float y, float z, float w)
{
  const float v[] = {x, y, z, w};
  ospSetParam(o, n, OSP_VEC4F, v);
}

void ospSetVec2i(OSPObject o, const char *n, int x, int y)
{
  const int v[] = {x, y};
  ospSetParam(o, n, OSP_VEC2I, v);
}

void ospSetVec3i(OSPObject o, const char *n, int x, int y, int z)
{
  const int v[] = {x, y, z};
  ospSetParam(o, n, OSP_VEC3I, v);
}

void ospSetVec4i(OSPObject o, const char *n, int x, int y, int z, int w)
{
  const int v[] = {x, y, z, w};
  ospSetParam(o, n, OSP_VEC4I, v);
}

void ospSetObjectAsData(OSPObject o, const char *n, OSPDataType type, OSPObject p)
{
  OSPData tmp = ospNewSharedData(&p, type, 1, 0, 1, 0, 1, 0, NULL, NULL);
  OSPData data = ospNewData(type, 1, 1, 1);
  ospCopyData(tmp, data, 0, 0, 0);
  ospSetParam(o, n, OSP_DATA, &data);
  ospRelease(tmp);
  ospRelease(data);
}

float ospRenderFrameBlocking(OSPFrameBuffer fb, OSPRenderer r, OSPCamera c, OSPWorld w)
{
  OSPFuture f = ospRenderFrame(fb, r, c, w);
  ospWait(f, OSP_TASK_FINISHED);
  ospRelease(f);
  return ospGetVariance(fb);
}

} // extern "C"
