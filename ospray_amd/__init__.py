"""MI355X-native OSPRay device: Python bindings of the C API, scenes."""
