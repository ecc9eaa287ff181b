// sky.cpp -- Preetham daylight baked into an equirectangular map (host side;
// the map then takes the HDRI path: importance table, NEE, MIS lookups).
// Binary64 arithmetic with the operation order of the reference's JavaScript
// (shared-samplers/src/preethamSky.ts, solarDiscCoverage.ts, hgPhase.ts), so
// each texel rounds to binary32 from the same value up to libm's last ulp.
#include "sky.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace vts {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 2 * kPi;
constexpr double kSolarRadius = 0.00436;  // ~0.25 degrees
constexpr double kHgLimit = 0.999999;

// Perez et al. luminance distribution F(theta, gamma) with coefficients A..E
// (Preetham et al. 1999, eq. 3); cos(theta) floored at 1e-4.
double perez(double cos_theta, double gamma, const double k[5]) {
  const double c = std::max(cos_theta, 1e-4);
  const double cg = std::cos(gamma);
  return (1 + k[0] * std::exp(k[1] / c)) * (1 + k[2] * std::exp(k[3] * gamma) + k[4] * cg * cg);
}

// The distribution coefficients as linear functions of turbidity T
// (Preetham et al. 1999, appendix A.2): luminance Y, chromaticities x and y.
void coefficients_Y(double T, double k[5]) {
  k[0] = 0.17872 * T - 1.46303;
  k[1] = -0.3554 * T + 0.42749;
  k[2] = -0.02266 * T + 5.32505;
  k[3] = 0.12064 * T - 2.57705;
  k[4] = -0.06696 * T + 0.37027;
}
void coefficients_x(double T, double k[5]) {
  k[0] = -0.01925 * T - 0.25922;
  k[1] = -0.06651 * T + 0.00081;
  k[2] = -0.00041 * T + 0.21247;
  k[3] = -0.06409 * T - 0.89887;
  k[4] = -0.00325 * T + 0.04517;
}
void coefficients_y(double T, double k[5]) {
  k[0] = -0.01669 * T - 0.26078;
  k[1] = -0.09495 * T + 0.00921;
  k[2] = -0.00792 * T + 0.21023;
  k[3] = -0.04405 * T - 1.65369;
  k[4] = -0.01092 * T + 0.05291;
}

// Zenith luminance and chromaticity (Preetham et al. 1999, appendix A.2).
double zenith_Y(double T, double theta_sun) {
  const double chi = (4.0 / 9.0 - T / 120.0) * (kPi - 2 * theta_sun);
  return (4.0453 * T - 4.971) * std::tan(chi) - 0.2155 * T + 2.4192;
}
void zenith_xy(double T, double ts, double* xz, double* yz) {
  const double T2 = T * T, ts2 = ts * ts, ts3 = ts2 * ts;
  *xz = (0.00166 * ts3 - 0.00375 * ts2 + 0.00209 * ts) * T2 +
        (-0.02903 * ts3 + 0.06377 * ts2 - 0.03202 * ts + 0.00394) * T +
        (0.11693 * ts3 - 0.21196 * ts2 + 0.06052 * ts + 0.25886);
  *yz = (0.00275 * ts3 - 0.0061 * ts2 + 0.00317 * ts) * T2 +
        (-0.04214 * ts3 + 0.0897 * ts2 - 0.04153 * ts + 0.00516) * T +
        (0.15346 * ts3 - 0.26756 * ts2 + 0.0667 * ts + 0.26688);
}

// CIE xyY to linear Rec.709 rgb (the sRGB XYZ -> RGB matrix), negatives
// clamped to zero.
void xyY_to_rgb(double x, double y, double Y, double rgb[3]) {
  rgb[0] = rgb[1] = rgb[2] = 0;
  if (y < 1e-10 || Y < 0) return;
  const double X = (Y / y) * x, Z = (Y / y) * (1 - x - y);
  rgb[0] = std::max(0.0, 3.2406 * X - 1.5372 * Y - 0.4986 * Z);
  rgb[1] = std::max(0.0, -0.9689 * X + 1.8758 * Y + 0.0415 * Z);
  rgb[2] = std::max(0.0, 0.0557 * X - 0.204 * Y + 1.057 * Z);
}

// Henyey-Greenstein phase density [1/sr], evaluated about the signed lobe
// axis (hgPhase.ts:61-77).
double henyey_greenstein(double cos_theta, double g) {
  cos_theta = std::max(-1.0, std::min(1.0, cos_theta));
  g = std::max(-kHgLimit, std::min(kHgLimit, g));
  const double a = std::fabs(g);
  const double c = g >= 0 ? cos_theta : -cos_theta;
  const double om = 1 - a;
  const double den = om * om + 2 * a * (1 - c);
  return (1 / (4 * kPi)) * (om * (1 + a)) / (den * std::sqrt(den));
}

double wrap_longitude(double phi) {
  const double w = std::fmod(phi, kTwoPi);
  return w < 0 ? w + kTwoPi : w;
}

bool finite(double v) { return std::isfinite(v); }

}  // namespace

std::string solar_coverage(uint32_t W, uint32_t H, const double sun[3], std::vector<double>& cov) {
  const double len = std::hypot(sun[0], sun[1], sun[2]);
  if (!finite(len) || len <= 0) return "the sun direction must be a finite non-zero vector";
  const double sx = sun[0] / len, sy = sun[1] / len, sz = sun[2] / len;
  const double theta_sun = std::acos(std::max(-1.0, std::min(1.0, sy)));
  const double sin_sun = std::hypot(sx, sz);
  // texture longitude u in [0, 1) is world phi + pi
  const double sun_phi = wrap_longitude(std::atan2(sz, sx) + kPi);
  const double cos_r = std::cos(kSolarRadius);
  const double cap = kTwoPi * (1 - cos_r);
  const double dth = kPi / H, dph = kTwoPi / W;
  cov.assign((size_t)W * H, 0.0);
  const double th_min = std::max(0.0, theta_sun - kSolarRadius);
  const double th_max = std::min(kPi, theta_sun + kSolarRadius);
  const int64_t row0 = std::max<int64_t>(0, (int64_t)std::floor(th_min / dth));
  const int64_t row1 = std::min<int64_t>((int64_t)H - 1, (int64_t)std::ceil(th_max / dth) - 1);
  double deposited = 0;
  if (sin_sun <= 1e-12) {
    // the cap around a pole: whole rows, closed form
    const bool north = sy >= 0;
    const double lo_cap = north ? 0.0 : kPi - kSolarRadius, hi_cap = north ? kSolarRadius : kPi;
    for (int64_t r = row0; r <= row1; ++r) {
      const double lo = std::max(r * dth, lo_cap), hi = std::min((r + 1) * dth, hi_cap);
      if (hi <= lo) continue;
      const double om = dph * (std::cos(lo) - std::cos(hi));
      for (uint32_t x = 0; x < W; ++x) cov[(size_t)r * W + x] = om;
      deposited += om * W;
    }
  } else {
    // deposit weight w over longitudes [a, b) of row r, texel by texel
    auto interval = [&](int64_t r, double a, double b, double w) {
      if (b <= a) return;
      const int64_t c0 = std::max<int64_t>(0, (int64_t)std::floor(a / dph));
      const int64_t c1 = std::min<int64_t>((int64_t)W - 1, (int64_t)std::ceil(b / dph) - 1);
      for (int64_t x = c0; x <= c1; ++x) {
        const double ov = std::max(0.0, std::min(b, (x + 1) * dph) - std::max(a, x * dph));
        if (ov <= 0) continue;
        cov[(size_t)r * W + x] += w * ov;
        deposited += w * ov;
      }
    };
    auto span = [&](int64_t r, double half, double w) {
      if (half >= kPi) {
        interval(r, 0, kTwoPi, w);
        return;
      }
      const double s = wrap_longitude(sun_phi - half), e = s + 2 * half;
      interval(r, s, std::min(e, kTwoPi), w);
      if (e > kTwoPi) interval(r, 0, e - kTwoPi, w);
    };
    const double target = kSolarRadius / 64;
    for (int64_t r = row0; r <= row1; ++r) {
      const double lo = std::max(r * dth, th_min), hi = std::min((r + 1) * dth, th_max);
      if (hi <= lo) continue;
      const double extent = hi - lo;
      // composite midpoint rule in theta, 16..512 steps
      const int64_t steps = std::max<int64_t>(16, std::min<int64_t>(512, (int64_t)std::ceil(extent / target)));
      const double step = extent / (double)steps;
      for (int64_t k = 0; k < steps; ++k) {
        const double th = lo + (k + 0.5) * step;
        const double st = std::sin(th);
        const double w = st * step;
        const double horiz = st * sin_sun;
        if (horiz <= DBL_EPSILON) {
          if (std::cos(th) * sy >= cos_r) span(r, kPi, w);
          continue;
        }
        // longitudes within the cap: cos(dphi) >= (cos r - cos th sy) / horiz
        const double t = (cos_r - std::cos(th) * sy) / horiz;
        if (t <= -1)
          span(r, kPi, w);
        else if (t < 1)
          span(r, std::acos(std::max(-1.0, std::min(1.0, t))), w);
      }
    }
  }
  if (!(deposited > 0) || !finite(deposited)) return "the solar cap could not be integrated";
  // conserve the analytic cap solid angle exactly, then express per texel
  // against the centre-sampled measure the bake integrates with
  const double norm = cap / deposited;
  for (int64_t r = row0; r <= row1; ++r) {
    const double centre = ((r + 0.5) / H) * kPi;
    const double measure = dth * dph * std::max(std::sin(centre), 1e-6);
    for (uint32_t x = 0; x < W; ++x) {
      double& c = cov[(size_t)r * W + x];
      c = (c * norm) / measure;
    }
  }
  return {};
}

std::string bake_preetham(const SkyParams& p, uint32_t W, uint32_t H, std::vector<float>& rgb, double* integral) {
  if (W < 1 || W > 32768 || H < 1 || H > 32768) return "the sky map's width and height must be in [1, 32768]";
  if ((uint64_t)W * H * 4 > 0x7fffffffull) return "the sky map is too large";
  for (int k = 0; k < 3; ++k)
    if (!finite(p.sun[k])) return "procedural-sky sunDirection must be finite";
  if (!finite(p.turbidity) || !finite(p.rayleigh) || !finite(p.mie_coefficient) || !finite(p.mie_g) ||
      !finite(p.intensity))
    return "procedural-sky turbidity, rayleigh, mieCoefficient, mieDirectionalG and intensity must be finite";
  double sun[3];
  {
    const double len = std::hypot(p.sun[0], p.sun[1], p.sun[2]);
    if (len < 1e-8) {
      sun[0] = 0, sun[1] = 1, sun[2] = 0;
    } else {
      for (int k = 0; k < 3; ++k) sun[k] = p.sun[k] / len;
    }
  }
  const double T = std::max(1.5, std::min(30.0, p.turbidity));
  const double rayleigh = std::max(0.0, p.rayleigh);
  const double mie = std::max(0.0, p.mie_coefficient);
  const double mie_scale = mie * 200;
  if (!finite(mie_scale)) return "procedural-sky mieCoefficient is too large";
  const double g = std::max(-kHgLimit, std::min(kHgLimit, p.mie_g));
  const double intensity = std::max(0.0, p.intensity);

  const double theta_sun = std::acos(std::max(-1.0, std::min(1.0, sun[1])));
  double kY[5], kx[5], ky[5];
  coefficients_Y(T, kY);
  coefficients_x(T, kx);
  coefficients_y(T, ky);
  const double zY = std::max(1e-5, zenith_Y(T, theta_sun)) * rayleigh * intensity;
  if (!finite(zY)) return "procedural-sky zenith luminance overflowed";
  double xz, yz;
  zenith_xy(T, theta_sun, &xz, &yz);
  const double nY = perez(1, theta_sun, kY), nx = perez(1, theta_sun, kx), ny = perez(1, theta_sun, ky);
  std::vector<double> cov;
  const std::string ce = solar_coverage(W, H, sun, cov);
  if (!ce.empty()) return ce;

  rgb.assign((size_t)W * H * 3, 0.0f);
  const double dph = kTwoPi / W;
  double total = 0;
  for (uint32_t py = 0; py < H; ++py) {
    const double theta = ((py + 0.5) / H) * kPi;
    const double ct = std::cos(theta), st = std::sin(theta);
    const double omega = dph * (std::cos((py / (double)H) * kPi) - std::cos(((py + 1) / (double)H) * kPi));
    for (uint32_t px = 0; px < W; ++px) {
      // u = phi / 2 pi + 0.5: column 0 is centred near -pi, +X at the middle
      const double phi = ((px + 0.5) / W - 0.5) * kTwoPi;
      const double dx = st * std::cos(phi), dy = ct, dz = st * std::sin(phi);
      const double cg = std::max(-1.0, std::min(1.0, dx * sun[0] + dy * sun[1] + dz * sun[2]));
      const double gamma = std::acos(cg);
      double c[3] = {0, 0, 0};
      if (ct >= -0.05) {
        const double sc = std::max(ct, 1e-4);
        const double sY = nY > 1e-12 ? zY * (perez(sc, gamma, kY) / nY) : 0;
        const double sx = nx > 1e-12 ? xz * (perez(sc, gamma, kx) / nx) : xz;
        const double sy = ny > 1e-12 ? yz * (perez(sc, gamma, ky) / ny) : yz;
        const double fade = std::min(1.0, (ct + 0.05) / 0.05);
        const double lum = std::max(0.0, sY) * fade;
        const double mie_y = mie_scale * henyey_greenstein(cg, g) * 4 * kPi * zY * fade;
        const double sun_y = cov[(size_t)py * W + px] * 500 * zY;
        const double Y = lum + mie_y + sun_y;
        // the sky's chromaticity, or the zenith's where only sun and haze shine
        const bool sky = lum > 1e-12 || Y < 1e-12;
        xyY_to_rgb(sky ? sx : xz, sky ? sy : yz, Y, c);
      }
      float* q = &rgb[((size_t)py * W + px) * 3];
      for (int k = 0; k < 3; ++k) {
        if (!finite(c[k])) return "procedural-sky radiance must be finite";
        q[k] = (float)c[k];
        if (!std::isfinite(q[k])) return "procedural-sky radiance must be representable in binary32";
      }
      const double w = std::max(0.0, (0.2126 * q[0] + 0.7152 * (double)q[1] + 0.0722 * (double)q[2]) * omega);
      total += w;
      if (!finite(total)) return "procedural-sky luminance integral overflowed";
    }
  }
  if (intensity > 0 && rayleigh > 0 && total == 0)
    return "positive procedural-sky radiance underflows entirely to zero in binary32";
  *integral = total;
  return {};
}

}  // namespace vts
