// glm_lite.h -- the few glm value types the reference's FBP3D interface
// uses in its signatures (glm::vec3 in render_sphere / render_ellipsoid,
// glm::ivec2 / ivec3 from get_size(); Source/FilteredBackProjection/FBP3D.h:26-27).
// Define CTRECON_USE_SYSTEM_GLM to take them from a real glm instead.
#pragma once

#ifdef CTRECON_USE_SYSTEM_GLM
#include <glm/glm.hpp>
#else
namespace glm {

struct vec2 {
    float x = 0, y = 0;
    vec2() = default;
    vec2(float a, float b) : x(a), y(b) {}
};
struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(float a) : x(a), y(a), z(a) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    vec4() = default;
    vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};
struct ivec2 {
    int x = 0, y = 0;
    ivec2() = default;
    ivec2(int a, int b) : x(a), y(b) {}
    bool operator==(const ivec2 &o) const { return x == o.x && y == o.y; }
    bool operator!=(const ivec2 &o) const { return !(*this == o); }
};
struct ivec3 {
    int x = 0, y = 0, z = 0;
    ivec3() = default;
    ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
    bool operator==(const ivec3 &o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const ivec3 &o) const { return !(*this == o); }
};

inline vec3 operator/(vec3 a, float s) { return vec3(a.x / s, a.y / s, a.z / s); }
inline vec3 operator*(vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }

}  // namespace glm
#endif
