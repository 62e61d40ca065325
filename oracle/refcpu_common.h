// ORACLE — test infrastructure only. Never linked into or called by the product library.
//
// Small FP64 helpers for the CPU restatement of the reference's coupled LB/DEM step.
// The vector type reproduces the operation order of the reference's tinyVector
// (myvector.cpp:48-152) so that results are bit-identical when compiled without
// floating-point contraction (-ffp-contract=off, no FMA target).
#pragma once
#include <cmath>
#include <cstdint>

namespace refcpu {

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    V3 operator/(double s) const { return V3(x / s, y / s, z / s); }
    V3& operator/=(double s) { x /= s; y /= s; z /= s; return *this; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    double norm2() const { return x * x + y * y + z * z; }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    void reset() { x = y = z = 0.0; }
    // tinyVector::insideSphere (myvector.cpp:171-176)
    bool insideSphere(const V3& c, double r) const { V3 d = *this - c; return d.norm2() < r * r; }
    // tinyVector::insidePlane (myvector.cpp:222-225)
    bool insidePlane(const V3& p, const V3& n) const { V3 d = *this - p; return n.dot(d) < 0.0; }
};
// operator*(double, tVect) (myvector.cpp:108-113): components times scalar
inline V3 operator*(double s, const V3& v) { return V3(v.x * s, v.y * s, v.z * s); }

// D3Q19 lattice (lattice.h:33-57, 134, 147-150)
constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// slip directions (lattice.h:137-140)
constexpr int SLIP1_CHECK[Q] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 3, 4, 5, 6, 1, 2, 6, 5};
constexpr int SLIP1[Q] = {0, 0, 0, 0, 0, 0, 0, 9, 10, 8, 7, 13, 14, 12, 11, 18, 17, 15, 16};
constexpr int SLIP2_CHECK[Q] = {0, 0, 0, 0, 0, 0, 0, 3, 4, 2, 1, 5, 6, 4, 3, 5, 6, 1, 2};
constexpr int SLIP2[Q] = {0, 0, 0, 0, 0, 0, 0, 10, 9, 7, 8, 14, 13, 11, 12, 17, 18, 16, 15};
inline V3 vdir(int j) { return V3(double(CX[j]), double(CY[j]), double(CZ[j])); }
inline double coeff(int j) { return j == 0 ? 12.0 / 36.0 : (j < 7 ? 2.0 / 36.0 : 1.0 / 36.0); }

// DEM.cpp:4-11 macros
inline int n_int(double x) { return x >= 0 ? (int)(x + 0.5) : (int)(x - 0.5); }
inline int s_int(double x) { return x >= 0 ? (int)(x) : (int)(x - 1.0); }
inline double dbox(double a, double b) { return a - s_int(a / b) * b; }
inline double n_image(double a, double b) { return a - n_int(a / b) * b; }
inline double dmax(double a, double b) { return a > b ? a : b; }

}  // namespace refcpu
