// Oracle vector algebra: a host restatement of the reference's src/type.h
// helpers, used only by the FP64 oracle (test infrastructure; never linked
// into libnbc). Every helper evaluates in the same order as the reference so
// that, compiled with -ffp-contract=off, results are reproducible bit for bit.
#pragma once
#include <cmath>

namespace orc {

struct d3 { double x, y, z; };
struct d4 { double x, y, z, w; };

inline d3 mk3(double x, double y, double z) { return d3{x, y, z}; }

// src/type.h:20-60
inline d3 operator+(const d3& a, const d3& b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator-(const d3& a, const d3& b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator*(double s, const d3& u) { return d3{s * u.x, s * u.y, s * u.z}; }
// src/type.h:148-190
inline d4 operator+(const d4& a, const d4& b) { return d4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline d4 operator*(double s, const d4& u) { return d4{s * u.x, s * u.y, s * u.z, s * u.w}; }

// src/type.h:72-89 (zero below 1e-20)
inline d3 Norm(const d3& v) {
    double t = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (t > 1e-20) return d3{v.x / t, v.y / t, v.z / t};
    return d3{0.0, 0.0, 0.0};
}
// src/type.h:91-110
inline double Dot(const d3& a, const d3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 Cross(const d3& a, const d3& b) {
    double x = a.y * b.z - a.z * b.y;
    double y = a.z * b.x - a.x * b.z;
    double z = a.x * b.y - a.y * b.x;
    return d3{x, y, z};
}
// src/type.h:112-126 (zero at or below |v|^2 = 1e-32)
inline double Length(const d3& v) {
    double t = v.x * v.x + v.y * v.y + v.z * v.z;
    return t > 1e-32 ? std::sqrt(t) : 0.0;
}
// src/type.h:223-236 (world -> body, "Cnb")
inline void Quat2DCM(const d4& q, double m[3][3]) {
    m[0][0] = q.x * q.x - q.y * q.y - q.z * q.z + q.w * q.w;
    m[1][0] = 2.0 * (q.x * q.y - q.z * q.w);
    m[2][0] = 2.0 * (q.x * q.z + q.y * q.w);
    m[0][1] = 2.0 * (q.x * q.y + q.z * q.w);
    m[1][1] = -q.x * q.x + q.y * q.y - q.z * q.z + q.w * q.w;
    m[2][1] = 2.0 * (q.y * q.z - q.x * q.w);
    m[0][2] = 2.0 * (q.x * q.z - q.y * q.w);
    m[1][2] = 2.0 * (q.y * q.z + q.x * q.w);
    m[2][2] = -q.x * q.x - q.y * q.y + q.z * q.z + q.w * q.w;
}
// src/type.h:238-250 (body -> world, "Cbn")
inline void Quat2IvDCM(const d4& q, double m[3][3]) {
    m[0][0] = q.x * q.x - q.y * q.y - q.z * q.z + q.w * q.w;
    m[1][0] = 2.0 * (q.x * q.y + q.z * q.w);
    m[2][0] = 2.0 * (q.x * q.z - q.y * q.w);
    m[0][1] = 2.0 * (q.x * q.y - q.z * q.w);
    m[1][1] = -q.x * q.x + q.y * q.y - q.z * q.z + q.w * q.w;
    m[2][1] = 2.0 * (q.y * q.z + q.x * q.w);
    m[0][2] = 2.0 * (q.x * q.z + q.y * q.w);
    m[1][2] = 2.0 * (q.y * q.z - q.x * q.w);
    m[2][2] = -q.x * q.x - q.y * q.y + q.z * q.z + q.w * q.w;
}
// src/type.h:296-305
inline d3 Map(const double m[3][3], const d3& u) {
    double a = m[0][0] * u.x + m[0][1] * u.y + m[0][2] * u.z;
    double b = m[1][0] * u.x + m[1][1] * u.y + m[1][2] * u.z;
    double c = m[2][0] * u.x + m[2][1] * u.y + m[2][2] * u.z;
    return d3{a, b, c};
}

inline bool finite3(const d3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
inline bool finite4(const d4& v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z) && std::isfinite(v.w);
}

}  // namespace orc
