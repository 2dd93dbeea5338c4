"""MI355X-native texture-space (Mesh Colors) path tracer."""
