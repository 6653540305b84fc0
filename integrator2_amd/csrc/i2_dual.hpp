// i2_dual.hpp -- forward-mode dual numbers (value, derivative along one
// direction) for the per-task part of the adjoint: panel geometry, the
// singular-part contexts and the closed-form singular integrals are
// differentiated along one vertex coordinate per lane, with the generic
// device functions of i2_device.hpp instantiated on Dual.  Branch tests use
// val(), so the branches taken are those of the FP64 forward pass.
#pragma once
#include "i2_device.hpp"

namespace i2ad {

struct Dual {
    double v, d;
    I2_HD Dual() : v(0.0), d(0.0) {}
    I2_HD Dual(double x) : v(x), d(0.0) {}
    I2_HD Dual(double x, double dx) : v(x), d(dx) {}
};

I2_HD __forceinline__ double val(const Dual& a) { return a.v; }
I2_HD __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
I2_HD __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
I2_HD __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
I2_HD __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
I2_HD __forceinline__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
I2_HD __forceinline__ Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
I2_HD __forceinline__ Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }

I2_HD __forceinline__ Dual sqrt(Dual a) { const double r = ::sqrt(a.v); return {r, 0.5 * a.d / r}; }
I2_HD __forceinline__ Dual log(Dual a) { return {::log(a.v), a.d / a.v}; }
I2_HD __forceinline__ Dual sin(Dual a) { return {::sin(a.v), ::cos(a.v) * a.d}; }
I2_HD __forceinline__ Dual cos(Dual a) { return {::cos(a.v), -::sin(a.v) * a.d}; }
I2_HD __forceinline__ Dual tan(Dual a) {
    const double t = ::tan(a.v);
    return {t, (1.0 + t * t) * a.d};
}
I2_HD __forceinline__ void sincos(Dual a, Dual* s, Dual* c) {
    double sv, cv;
    ::sincos(a.v, &sv, &cv);
    *s = {sv, cv * a.d};
    *c = {cv, -sv * a.d};
}
I2_HD __forceinline__ Dual acos(Dual a) { return {::acos(a.v), -a.d / ::sqrt(1.0 - a.v * a.v)}; }
I2_HD __forceinline__ Dual atan2(Dual y, Dual x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return {::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}

using D3 = i2d::v3t<Dual>;
using Q4 = i2d::q4t<Dual>;

}  // namespace i2ad
