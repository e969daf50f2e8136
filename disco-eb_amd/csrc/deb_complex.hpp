// deb_complex.hpp -- complex arithmetic on double2 and the complex log-gamma, shared by the
// line-of-sight coefficients (deb_cmb.hip) and the FFTLog correlation function (deb_lss.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace deb {
namespace cplx {

constexpr double kPiC = 3.141592653589793;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 clog(double2 a) { return make_double2(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }

// log Gamma(z): Lanczos approximation, g = 7, 9 terms, reflection for Re z <= 1/2 (util.py:13-45)
__device__ inline double2 lngamma_complex(double2 z) {
  const double c[9] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                       -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                       -176.61502916214059906584551354, 12.507343278686904814458936853,
                       -0.13857109526572011689554707, 9.984369578019570859563e-6, 1.50563273514931155834e-7};
  const bool refl = z.x <= 0.5;
  double2 zz = refl ? make_double2(1.0 - z.x, -z.y) : z;
  zz.x -= 1.0;
  double2 Ag = make_double2(c[0], 0.0);
  for (int j = 1; j < 9; ++j) {
    const double2 t = make_double2(zz.x + j, zz.y);
    const double n2 = t.x * t.x + t.y * t.y;
    Ag.x += c[j] / n2 * t.x;
    Ag.y += c[j] / n2 * (-t.y);
  }
  const double2 zh = make_double2(zz.x + 0.5, zz.y), z75 = make_double2(zz.x + 7.5, zz.y);
  double2 r = cmul(zh, clog(z75));
  r = csub(r, z75);
  r.x += 0.9189385332046727418;
  r = cadd(r, clog(Ag));
  if (!refl) return r;
  // ln pi - ln sin(pi z) - lnGamma(1 - z); sin(pi (x + i y)) = sin(pi x) cosh(pi y) + i cos(pi x) sinh(pi y)
  double sx, cx;
  sincospi(z.x, &sx, &cx);
  const double2 sn = make_double2(sx * cosh(kPiC * z.y), cx * sinh(kPiC * z.y));
  double2 out = make_double2(1.14472988584940017414342735135, 0.0);
  out = csub(out, clog(sn));
  return csub(out, r);
}


}  // namespace cplx
}  // namespace deb
