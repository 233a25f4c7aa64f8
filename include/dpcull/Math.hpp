// Math.hpp -- the value types the host interfaces exchange.  Same memory layout
// as nvpro-pipeline's dp::math types: Vec3f/Vec4f are 3/4 packed floats,
// Mat44f is 16 floats, row-major, row-vector convention (dp/math/Vecnt.h:200,
// dp/math/Matmnt.h:199), Box3f is (lower, upper) (dp/math/Boxnt.h:90), so a
// reference array of Mat44f can be handed over by pointer.
#pragma once

#include <cfloat>

namespace dpcull {

struct Vec3f {
  float v[3];
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};

struct Vec4f {
  float v[4];
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};

struct Mat44f {
  float m[16];
  float* operator[](int r) { return m + 4 * r; }
  const float* operator[](int r) const { return m + 4 * r; }
  static Mat44f identity() { return Mat44f{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}}; }
};

// A box stores lower and upper corners; an empty box is (FLT_MAX, -FLT_MAX).
struct Box3f {
  Vec3f lower{{FLT_MAX, FLT_MAX, FLT_MAX}};
  Vec3f upper{{-FLT_MAX, -FLT_MAX, -FLT_MAX}};
};

static_assert(sizeof(Vec3f) == 12 && sizeof(Vec4f) == 16 && sizeof(Mat44f) == 64 && sizeof(Box3f) == 24,
              "dpcull math types must match dp::math's layout");

}  // namespace dpcull
