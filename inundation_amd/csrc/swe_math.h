// Device arithmetic of one shallow-water edge and cell, FP32, for gfx950.
//
// Every function is a sequence of IEEE-754 single-precision operations
// (+, -, *, / and sqrt, all correctly rounded; the library is compiled with
// -ffp-contract=off so no multiply-add is fused behind our back). The order of
// operations inside each expression follows the reference GLSL, so results
// depend only on the inputs:
//   safe_vel         inundation/vulkan/shaders/flux.py:63-66
//   hllc_speeds      inundation/vulkan/shaders/flux.py:75-95
//   hllc             inundation/vulkan/shaders/flux.py:98-140
//   face_states      inundation/vulkan/shaders/flux.py:146-169
//   wave_bound       inundation/vulkan/shaders/flux.py:172-182
//   film_drag        inundation/vulkan/shaders/flux.py:46-51
//   two_sum / add_compensated   inundation/vulkan/shaders/flux.py:301-319
// GLSL min/max/clamp are spelled out as comparisons (GLSL defines
// min(x, y) = y < x ? y : x), so signed zeros come out the same everywhere.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace inu {

// Scalar parameters of a run, with the float products the reference forms in
// every call precomputed once (same single rounding, so the same bits).
struct Phys {
  float g;         // gravity
  float dry;       // dry_tol
  float cfl;       // CFL number
  float e_film;    // FILM_DEPTHS * dry_tol
  float e_film2;   // e_film * e_film
  float half_g;    // 0.5f * g   (== g * 0.5f)
  float sqrt_g_e;  // sqrt(g / e_film)
  float inv_dry;   // 1 / dry_tol
  float sqrt_g;    // sqrt(g)
  float adj_thin;  // adjoint only: face depth below which a side's velocity and celerity are held
};

__host__ __device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__host__ __device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__host__ __device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

// Square root, correctly rounded (the same bits as sqrtf). For 0 and for
// x >= 2^-96 the library's own sequence is the hardware estimate moved one
// ulp down or up by the signs of two fused residuals; its rescaling of tiny
// arguments and its special-case step do nothing there, so they are skipped
// unless the argument is below 2^-96 (or negative or NaN), which takes the
// library path. Infinity comes out of the residual steps unchanged.
__device__ __forceinline__ float sqrt_rn(float x) {
  if (!(x >= 0x1p-96f) && x != 0.0f) return sqrtf(x);
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  const float r = (fmaf(-dn, s, x) <= 0.0f) ? dn : s;
  return (fmaf(-up, s, x) > 0.0f) ? up : r;
}

__device__ __forceinline__ float safe_vel(const Phys& p, float h, float hq) {
  return (h > p.dry) ? (2.0f * h * hq) / (h * h + gmax(h * h, p.e_film2)) : 0.0f;
}

__device__ __forceinline__ void hllc_speeds(const Phys& p, float hL, float unL, float hR, float unR,
                                            float& SL, float& SR) {
  const float cL = sqrt_rn(p.g * gmax(hL, 0.0f));
  const float cR = sqrt_rn(p.g * gmax(hR, 0.0f));
  const float sqL = sqrt_rn(gmax(hL, 0.0f));
  const float sqR = sqrt_rn(gmax(hR, 0.0f));
  const float den = sqL + sqR;
  const float u_roe = (den > 1e-12f) ? (sqL * unL + sqR * unR) / den : 0.0f;
  const float c_roe = sqrt_rn(p.half_g * (hL + hR));
  float sl = gmin(unL - cL, u_roe - c_roe);
  float sr = gmax(unR + cR, u_roe + c_roe);
  const float wetL = gclamp(hL / p.dry - 1.0f, 0.0f, 1.0f);
  const float wetR = gclamp(hR / p.dry - 1.0f, 0.0f, 1.0f);
  const float dryLeft = (1.0f - wetL) * wetR;
  const float dryRight = (1.0f - wetR) * wetL;
  const float roe = 1.0f - dryLeft - dryRight;
  SL = roe * sl + dryLeft * (unR - 2.0f * cR) + dryRight * (unL - cL);
  SR = roe * sr + dryLeft * (unR + cR) + dryRight * (unL + 2.0f * cL);
}

__device__ __forceinline__ void hllc(const Phys& p, float hL, float hqnL, float hqtL, float hR,
                                     float hqnR, float hqtR, float& Fh, float& Fhn, float& Fht,
                                     float& Smax) {
  const float unL = safe_vel(p, hL, hqnL);
  const float utL = safe_vel(p, hL, hqtL);
  const float unR = safe_vel(p, hR, hqnR);
  const float utR = safe_vel(p, hR, hqtR);
  float SL, SR;
  hllc_speeds(p, hL, unL, hR, unR, SL, SR);
  Smax = gmax(fabsf(SL), fabsf(SR));

  const float num_sm = SL * hR * (unR - SR) - SR * hL * (unL - SL);
  const float den_sm = hR * (unR - SR) - hL * (unL - SL);
  const float SM = (fabsf(den_sm) > 1e-12f) ? num_sm / den_sm : 0.0f;

  const float FL_h = hL * unL;
  const float FL_hn = hL * unL * unL + p.half_g * hL * hL;
  const float FL_ht = hL * unL * utL;
  const float FR_h = hR * unR;
  const float FR_hn = hR * unR * unR + p.half_g * hR * hR;
  const float FR_ht = hR * unR * utR;

  // Region of the fan that holds the face (Toro 2001, eq. 10.26); only the
  // star state that is used is formed.
  if (SL >= 0.0f) {
    Fh = FL_h; Fhn = FL_hn; Fht = FL_ht;
  } else if (SM >= 0.0f) {
    const float dSL = SL - SM;
    const float kL = hL * (SL - unL) / ((fabsf(dSL) > 1e-14f) ? dSL : 1e-14f);
    Fh = FL_h + SL * (kL - hL);
    Fhn = FL_hn + SL * (kL * SM - hL * unL);
    Fht = FL_ht + SL * (kL * utL - hL * utL);
  } else if (SR >= 0.0f) {
    const float dSR = SR - SM;
    const float kR = hR * (SR - unR) / ((fabsf(dSR) > 1e-14f) ? dSR : 1e-14f);
    Fh = FR_h + SR * (kR - hR);
    Fhn = FR_hn + SR * (kR * SM - hR * unR);
    Fht = FR_ht + SR * (kR * utR - hR * utR);
  } else {
    Fh = FR_h; Fhn = FR_hn; Fht = FR_ht;
  }
}

__device__ __forceinline__ void face_states(const Phys& p, float hL, float huL, float hvL, float dzL,
                                            float hR, float huR, float hvR, float dzR, float nx,
                                            float ny, float& hLs, float& hqnL, float& hqtL,
                                            float& hRs, float& hqnR, float& hqtR) {
  hLs = gmax(0.0f, hL - dzL);
  hRs = gmax(0.0f, hR - dzR);
  const float uL = safe_vel(p, hL, huL);
  const float vL = safe_vel(p, hL, hvL);
  const float uR = safe_vel(p, hR, huR);
  const float vR = safe_vel(p, hR, hvR);
  const float huLs = hLs * uL;
  const float hvLs = hLs * vL;
  const float huRs = hRs * uR;
  const float hvRs = hRs * vR;
  hqnL = huLs * nx + hvLs * ny;
  hqtL = -huLs * ny + hvLs * nx;
  hqnR = huRs * nx + hvRs * ny;
  hqtR = -huRs * ny + hvRs * nx;
}

__device__ __forceinline__ float wave_bound(const Phys& p, float hL, float huL, float hvL, float dzL,
                                            float hR, float huR, float hvR, float dzR, float nx,
                                            float ny) {
  float hLs, hqnL, hqtL, hRs, hqnR, hqtR;
  face_states(p, hL, huL, hvL, dzL, hR, huR, hvR, dzR, nx, ny, hLs, hqnL, hqtL, hRs, hqnR, hqtR);
  float SL, SR;
  hllc_speeds(p, hLs, safe_vel(p, hLs, hqnL), hRs, safe_vel(p, hRs, hqnR), SL, SR);
  return gmax(fabsf(SL), fabsf(SR));
}

// Reflective wall: the ghost mirrors the inner state with the normal momentum
// negated (reference resident.py:126-129).
__device__ __forceinline__ void wall_ghost(float huL, float hvL, float nx, float ny, float& huR,
                                           float& hvR) {
  huR = huL - 2.0f * (huL * nx + hvL * ny) * nx;
  hvR = hvL - 2.0f * (huL * nx + hvL * ny) * ny;
}

// The per-edge words of one live edge: mass flux, left momentum (x, y), right
// momentum (x, y), all times the edge length, and the CFL term Smax * len
// (reference resident.py:165-184 with flux.py:98-169, 271-272).
// fh, if given, receives the HLLC mass flux before a wall zeroes it.
__device__ __forceinline__ void edge_words(const Phys& p, bool wall, float hL, float huL, float hvL,
                                           float hR, float huR, float hvR, float dzL, float dzR,
                                           float nx, float ny, float len, float w[6], float* fh = nullptr) {
  float hLs, hqnL, hqtL, hRs, hqnR, hqtR;
  face_states(p, hL, huL, hvL, dzL, hR, huR, hvR, dzR, nx, ny, hLs, hqnL, hqtL, hRs, hqnR, hqtR);
  float Fh, Fhn, Fht, Smax;
  hllc(p, hLs, hqnL, hqtL, hRs, hqnR, hqtR, Fh, Fhn, Fht, Smax);
  if (fh) *fh = Fh;
  if (wall) Fh = 0.0f;  // a wall passes no mass
  const float Fhu = Fhn * nx - Fht * ny;
  const float Fhv = Fhn * ny + Fht * nx;
  const float SL_g = p.half_g * (hL * hL - hLs * hLs);
  const float SR_g = p.half_g * (hR * hR - hRs * hRs);
  w[0] = Fh * len;
  w[1] = Fhu * len + SL_g * nx * len;
  w[2] = Fhv * len + SL_g * ny * len;
  w[3] = Fhu * len + SR_g * nx * len;
  w[4] = Fhv * len + SR_g * ny * len;
  w[5] = Smax * len;
}

__device__ __forceinline__ float film_drag(const Phys& p, float h, float dt) {
  if (h >= p.e_film) return 1.0f;
  const float rate = (p.e_film2 - h * h) / (2.0f * h * h) * p.sqrt_g_e;
  return 1.0f / (1.0f + dt * rate);
}

// Knuth two-sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  const float t = a + b;
  const float bb = t - a;
  e = (a - (t - bb)) + (b - bb);
  s = t;
}

// (hi + lo) + inc as a renormalised pair, clamped at zero.
__device__ __forceinline__ void add_compensated(float hi, float lo, float inc, float& h_new,
                                                float& lo_new) {
  float s, e;
  two_sum(hi, inc, s, e);
  e += lo;
  two_sum(s, e, h_new, lo_new);
  if (h_new < 0.0f || (h_new == 0.0f && lo_new < 0.0f)) {
    h_new = 0.0f;
    lo_new = 0.0f;
  }
}

// floor(e / 3) for an exponent e in [-149, 128], branch-free: e + 387 is
// positive and below 2^17 / 3, where (x * 43691) >> 17 is floor(x / 3).
__host__ __device__ __forceinline__ int floor_div3(int e) {
  return (int)(((uint32_t)(e + 387) * 43691u) >> 17) - 129;
}

// x^(1/3) for positive normal x, from exact exponent splitting and two Halley
// steps on the mantissa in [1, 8). Only IEEE operations, so the result is the
// same on every conforming device; within 2 ulp of the true cube root.
__device__ __forceinline__ float cbrt_ieee(float x) {
  const uint32_t u = __float_as_uint(x);
  const int e = (int)((u >> 23) & 0xffu) - 127;
  const int k = floor_div3(e);
  const int r = e - 3 * k;  // 0, 1 or 2
  const float m = __uint_as_float((u & 0x7fffffu) | ((uint32_t)(127 + r) << 23));
  float y = 0.80169147f + m * (0.24778609f - m * 0.0127262f);  // cbrt on [1, 8), 3.7 %
  float t = y * y * y;
  y = y * (t + 2.0f * m) / (2.0f * t + m);
  t = y * y * y;
  y = y * (t + 2.0f * m) / (2.0f * t + m);
  return y * __uint_as_float((uint32_t)(127 + k) << 23);
}

// h^(4/3) as h * cbrt(h): the Manning friction depth factor.
__device__ __forceinline__ float pow43(float h) { return h * cbrt_ieee(h); }

// ---------------------------------------------------------------------------
// Open boundaries (reference shaders/flux.py:194-282): the state beyond an open
// edge, from the cell inside it. Both arithmetic orders use these, in the
// reference's literal order; the two fractional powers are formed from the
// IEEE-only cube root (h^(2/3) = cbrt(h)^2, q^(1/3) = cbrt(q)).
// ---------------------------------------------------------------------------
constexpr uint32_t kOpenWall = 0u, kOpenFree = 1u, kOpenNormal = 2u, kOpenStage = 3u, kOpenInflow = 4u;
constexpr int kOpenNewtonSteps = 4;

// h^(2/3), 0 below the smallest normal float.
__device__ __forceinline__ float pow23(float h) {
  if (!(h >= 1.17549435e-38f)) return 0.0f;
  const float c = cbrt_ieee(h);
  return c * c;
}

// Depth beyond an inflow edge of discharge q (> 0) per unit width: the root of
// -q/hG + 2 sqrt(g hG) = riemann, the outgoing characteristic, by Newton in
// s = sqrt(hG) from below, each step at least halving s.
__device__ __forceinline__ float inflow_depth(const Phys& p, float q, float riemann) {
  const float root_g = p.sqrt_g;
  float s = gmax(cbrt_ieee(q / (2.0f * root_g)), riemann / (2.0f * root_g));
  for (int k = 0; k < kOpenNewtonSteps; ++k) {
    const float f = 2.0f * root_g * s - q / (s * s) - riemann;
    const float df = 2.0f * root_g + 2.0f * q / (s * s * s);
    s = gmax(s - f / df, 0.5f * s);
  }
  return s * s;
}

// Ghost state beyond an open edge of condition `kind`. `value` is sqrt(S)/n
// for normal depth, the depth over the cell's bed for a stage, the discharge
// per unit width for an inflow; wall and free outflow take none.
__device__ __forceinline__ void ghost_state(const Phys& p, uint32_t kind, float value, float hL, float huL, float hvL,
                                            float nx, float ny, float& hG, float& huG, float& hvG) {
  const float uL = safe_vel(p, hL, huL);
  const float vL = safe_vel(p, hL, hvL);
  const float unL = uL * nx + vL * ny;
  float ut = -uL * ny + vL * nx;
  const float cL = sqrt_rn(p.g * gmax(hL, 0.0f));
  float un = -unL;
  hG = hL;
  if (kind == kOpenFree) {
    un = fabsf(unL);
  } else if (kind == kOpenNormal) {
    un = value * pow23(gmax(hL, 0.0f));
  } else if (kind == kOpenStage) {
    // A level cannot reach a boundary against supercritical outflow.
    if (hL < p.dry || unL < cL) {
      hG = value;
      un = unL + 2.0f * (cL - sqrt_rn(p.g * value));
    } else {
      un = unL;
    }
  } else if (kind == kOpenInflow && value > 0.0f) {
    hG = inflow_depth(p, value, unL + 2.0f * cL);
    un = -value / hG;
    ut = 0.0f;
  }
  huG = hG * (un * nx - ut * ny);
  hvG = hG * (un * ny + ut * nx);
}

// An inflow keeps its edge live while dry, or its volume would be lost.
__device__ __forceinline__ bool open_forced(uint32_t kind, float value) { return kind == kOpenInflow && value > 0.0f; }

// Free outflow and normal depth never draw water in.
__device__ __forceinline__ bool open_outflow_only(uint32_t kind) { return kind == kOpenFree || kind == kOpenNormal; }

// Mass flux through an open edge whose HLLC mass flux is fh: outflow
// conditions pass none inward, a stage passes fh, an inflow exactly its
// discharge, a closed edge nothing.
__device__ __forceinline__ float boundary_mass(uint32_t kind, float value, float fh) {
  if (kind == kOpenFree || kind == kOpenNormal) return gmax(fh, 0.0f);
  if (kind == kOpenStage) return fh;
  if (kind == kOpenInflow) return -value;
  return 0.0f;
}

// ---------------------------------------------------------------------------
// Fast arithmetic: the same scheme, algebraically unchanged, in fewer
// instructions.
//
//  - Per cell, once: the desingularised velocity (u, v) = (hu, hv) vf(h) and
//    the face factor phi(h) = h vf(h), where vf(h) = 2h / (h^2 + max(h^2,
//    e^2)) is 1/h for h >= e_film, so phi is exactly 1 there. A side's face
//    velocity is its cell velocity, turned into the edge frame, times
//    phi(face depth): the reference's safe_vel(h*, h* u) (flux.py:63-66,
//    146-169). Only a face that steps up has its own phi, and only a thin
//    one (below e_film) needs a division for it.
//  - A wall's mirror state is the inside velocity with its normal part
//    negated (resident.py:126-129), taken in the edge frame directly.
//  - Divisions are products with a correctly rounded reciprocal, rcp(x)
//    (below), whose divisors stay in [2^-126, 2^126].
//  - min, max and the [0, 1] clamp are the hardware's (minNum / maxNum with
//    -0 below +0; the clamp takes NaN and -0 to +0).
//  - c = sqrt(g) * sqrt(h) instead of sqrt(g * h); (h / dry_tol) as h *
//    (1 / dry_tol); one star state (the side the face lies on) formed
//    branch-free; a * b + c as fmaf(a, b, c) where the exact order adds a
//    product; h^(-4/3) for Manning from a division-free inverse cube root.
// Rounding differs from the exact order by a few ulp; results depend only on
// the inputs (IEEE operations, fmaf and the reciprocal, no contraction), and
// the FP32 oracle restates the same sequence.
// ---------------------------------------------------------------------------
namespace fast {

// The largest divisor rcp takes exactly.
constexpr float kRcpMax = 0x1p126f;

// 1/x, correctly rounded, for 2^-126 <= |x| <= 2^126: the hardware estimate
// and one Newton step with a fused residual. Equal to the IEEE division 1/x
// for every float in that range (checked exhaustively, inu_probe_rcp), and
// not outside it (zero, subnormals, |x| > 2^126, infinities).
__device__ __forceinline__ float rcp(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, y, 1.0f), y, y);
}

// x clamped to [-2^126, 2^126] (NaN to -2^126): a divisor already bounded
// away from zero brought inside rcp's range.
__device__ __forceinline__ float rcp_range(float x) { return __builtin_amdgcn_fmed3f(x, -kRcpMax, kRcpMax); }

// The hardware's min and max (a NaN operand yields the other; -0 < +0) and
// [0, 1] clamp (NaN and -0 give +0).
__device__ __forceinline__ float vmin(float x, float y) { return __builtin_fminf(x, y); }
__device__ __forceinline__ float vmax(float x, float y) { return __builtin_fmaxf(x, y); }
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// Face factor of a thin depth h < e_film: 2h^2 / (h^2 + e^2) above dry_tol,
// else 0. (h^2 + e^2 lies in [e^2, 2e^2], inside rcp's range for the dry_tol
// inu_set_params accepts.)
__device__ __forceinline__ float thin_phi(const Phys& p, float h) {
  const float hh = h * h;
  return (h > p.dry) ? (2.0f * hh) * rcp(hh + p.e_film2) : 0.0f;
}

// vf(h) and phi(h) = h vf(h):
//   h >= e_film:      vf = 1 / min(h, 2^126),   phi = 1;
//   dry < h < e_film: vf = 2h / (h^2 + e^2),    phi = 2h^2 / (h^2 + e^2);
//   otherwise (and NaN): both 0.
__device__ __forceinline__ void factors(const Phys& p, float h, float& vf, float& phi) {
  const bool thick = h >= p.e_film;
  const float hh = h * h;
  const float r = rcp(thick ? vmin(h, kRcpMax) : hh + p.e_film2);
  vf = thick ? r : (2.0f * h) * r;
  phi = thick ? 1.0f : (2.0f * hh) * r;
  if (!(h > p.dry)) {
    vf = 0.0f;
    phi = 0.0f;
  }
}

// One side of an edge as the edge arithmetic reads it: depth, desingularised
// velocity, compensated low word (zero without compensation), phi(h) and
// sqrt(max(h, 0)).
struct Cell {
  float h, u, v, lo, phi, sq;
};

__device__ __forceinline__ Cell cell(const Phys& p, float h, float hu, float hv, float lo) {
  float vf, phi;
  factors(p, h, vf, phi);
  return Cell{h, hu * vf, hv * vf, lo, phi, sqrt_rn(vmax(h, 0.0f))};
}

// Face states of an edge with signed bed step d (d > 0: the left face steps
// up by d; d < 0: the right one by -d; at most one side steps). The stepped
// face depth is the other side's depth plus the surface difference from the
// full compensated depths (h + lo), so level water gives both sides the same
// bits: t = ((hL - hR) - d) + (loL - loR), hL* = max(hR + t, 0), hR* =
// max(hL - t, 0). `mirror`: R is L's wall mirror.
struct Face {
  float hL, hR, sqL, sqR, unL, utL, unR, utR;
};

__device__ __forceinline__ Face face(const Phys& p, const Cell& L, const Cell& R, bool mirror, float d, float nx,
                                     float ny) {
  Face f;
  const float t = ((L.h - R.h) - d) + (L.lo - R.lo);
  const bool stepL = d > 0.0f;
  const bool stepR = d < 0.0f;
  f.hL = stepL ? vmax(R.h + t, 0.0f) : L.h;
  f.hR = stepR ? vmax(L.h - t, 0.0f) : R.h;
  const float hs = stepL ? f.hL : f.hR;
  const float sqs = sqrt_rn(hs);
  float phs = 1.0f;
  if (!(hs >= p.e_film)) phs = thin_phi(p, hs);
  f.sqL = stepL ? sqs : L.sq;
  f.sqR = stepR ? sqs : R.sq;
  const float phL = stepL ? phs : L.phi;
  const float phR = stepR ? phs : R.phi;
  const float nL = fmaf(L.u, nx, L.v * ny);
  const float tL = fmaf(L.v, nx, -(L.u * ny));
  const float nR = mirror ? -nL : fmaf(R.u, nx, R.v * ny);
  const float tR = mirror ? tL : fmaf(R.v, nx, -(R.u * ny));
  f.unL = nL * phL;
  f.utL = tL * phL;
  f.unR = nR * phR;
  f.utR = tR * phR;
  return f;
}

// HLLC wave speeds (reference flux.py:75-95): Einfeldt/Roe bounds blended
// with Toro's dry-bed speeds over one dry_tol of face depth.
__device__ __forceinline__ void speeds(const Phys& p, const Face& f, float& SL, float& SR) {
  const float cL = p.sqrt_g * f.sqL;
  const float cR = p.sqrt_g * f.sqR;
  const float den = f.sqL + f.sqR;
  const float u_roe = (den > 1e-12f) ? fmaf(f.sqL, f.unL, f.sqR * f.unR) * rcp(vmin(den, kRcpMax)) : 0.0f;
  const float c_roe = sqrt_rn(p.half_g * (f.hL + f.hR));
  const float sl = vmin(f.unL - cL, u_roe - c_roe);
  const float sr = vmax(f.unR + cR, u_roe + c_roe);
  const float wetL = clamp01(fmaf(f.hL, p.inv_dry, -1.0f));
  const float wetR = clamp01(fmaf(f.hR, p.inv_dry, -1.0f));
  const float dryLeft = (1.0f - wetL) * wetR;
  const float dryRight = (1.0f - wetR) * wetL;
  const float roe = (1.0f - dryLeft) - dryRight;
  SL = fmaf(roe, sl, fmaf(dryLeft, f.unR - 2.0f * cR, dryRight * (f.unL - cL)));
  SR = fmaf(roe, sr, fmaf(dryLeft, f.unR + cR, dryRight * (f.unL + 2.0f * cL)));
}

// HLLC flux in the edge frame (reference flux.py:98-140).
__device__ __forceinline__ void hllc(const Phys& p, const Face& f, float& Fh, float& Fhn, float& Fht,
                                     float& Smax) {
  float SL, SR;
  speeds(p, f, SL, SR);
  Smax = vmax(fabsf(SL), fabsf(SR));
  const float hL = f.hL, hR = f.hR, unL = f.unL, unR = f.unR, utL = f.utL, utR = f.utR;
  // Toro's contact speed, numerator regrouped as SL SR (hL - hR) +
  // (SL hR unR - SR hL unL): exactly zero for still water.
  const float num_sm = fmaf(SL * SR, hL - hR, fmaf(SL * hR, unR, -(SR * hL * unL)));
  const float den_sm = fmaf(hR, unR - SR, -(hL * (unL - SL)));
  const float SM = (fabsf(den_sm) > 1e-12f) ? num_sm * rcp(rcp_range(den_sm)) : 0.0f;
  const float FL_h = hL * unL;
  const float FL_hn = fmaf(FL_h, unL, p.half_g * hL * hL);
  const float FL_ht = FL_h * utL;
  const float FR_h = hR * unR;
  const float FR_hn = fmaf(FR_h, unR, p.half_g * hR * hR);
  const float FR_ht = FR_h * utR;
  // The side whose state and flux the face takes: the left one when SL >= 0
  // (the face is left of the fan) or SM >= 0 (left of the contact), else
  // the right one. With SL < 0 this is the star state's side.
  const bool left = SM >= 0.0f || SL >= 0.0f;
  const float S = left ? SL : SR;
  const float hS = left ? hL : hR;
  const float unS = left ? unL : unR;
  const float utS = left ? utL : utR;
  const float FS_h = left ? FL_h : FR_h;
  const float FS_hn = left ? FL_hn : FR_hn;
  const float FS_ht = left ? FL_ht : FR_ht;
  // Star depth minus side depth, hS (SM - unS) / (S - SM): zero when the
  // contact does not move relative to the side. One reciprocal for both
  // cases: hS (SM - unS) / dS, or hS (S - unS) / 1e-14 - hS when the star
  // region is degenerate.
  const float dS = S - SM;
  const bool dS_ok = fabsf(dS) > 1e-14f;
  const float ks_q = (hS * (dS_ok ? SM - unS : S - unS)) * rcp(dS_ok ? rcp_range(dS) : 1e-14f);
  const float ks_h = dS_ok ? ks_q : ks_q - hS;
  const float kS = hS + ks_h;
  // Toro's four regions: F_L for SL >= 0, F_R for SR < 0 <= ... (right of
  // the fan), else the star flux of the chosen side.
  const bool outside = SL >= 0.0f || (!left && !(SR >= 0.0f));
  Fh = outside ? FS_h : fmaf(S, ks_h, FS_h);
  Fhn = outside ? FS_hn : fmaf(S, fmaf(kS, SM, -(hS * unS)), FS_hn);
  Fht = outside ? FS_ht : fmaf(S, ks_h * utS, FS_ht);
}

// The edge's six words (reference resident.py:165-184 with flux.py:98-169,
// 271-272): mass flux, left momentum (x, y), right momentum (x, y), all times
// the edge length, and the CFL term Smax * len. `wall` passes no mass; fh,
// if given, receives the HLLC mass flux before that.
__device__ __forceinline__ void edge_words(const Phys& p, bool wall, const Cell& L, const Cell& R, bool mirror,
                                           float d, float nx, float ny, float len, float w[6],
                                           float* fh = nullptr) {
  const Face f = face(p, L, R, mirror, d, nx, ny);
  float Fh, Fhn, Fht, Smax;
  hllc(p, f, Fh, Fhn, Fht, Smax);
  if (fh) *fh = Fh;
  if (wall) Fh = 0.0f;
  // Normal momentum flux plus the hydrostatic correction of each side,
  // grouped as (Fhn - P(face depth)) + P(cell depth): still water gives
  // exactly the cell's own pressure on every edge.
  const float nL = (Fhn - p.half_g * f.hL * f.hL) + p.half_g * L.h * L.h;
  const float nR = (Fhn - p.half_g * f.hR * f.hR) + p.half_g * R.h * R.h;
  const float tx = Fht * nx, ty = Fht * ny;
  w[0] = Fh * len;
  w[1] = fmaf(nL, nx, -ty) * len;
  w[2] = fmaf(nL, ny, tx) * len;
  w[3] = fmaf(nR, nx, -ty) * len;
  w[4] = fmaf(nR, ny, tx) * len;
  w[5] = Smax * len;
}

// The edge's words as the stage kernels keep them, one record per side:
// left (w0, w1, w2, w5) and right (-w0, -w3, -w4, w5), the negated words
// formed directly (a product with -len, the same bits as negating w).
__device__ __forceinline__ void edge_sides(const Phys& p, bool wall, const Cell& L, const Cell& R, bool mirror,
                                           float d, float nx, float ny, float len, float4& lw, float4& rw) {
  const Face f = face(p, L, R, mirror, d, nx, ny);
  float Fh, Fhn, Fht, Smax;
  hllc(p, f, Fh, Fhn, Fht, Smax);
  if (wall) Fh = 0.0f;
  const float nL = (Fhn - p.half_g * f.hL * f.hL) + p.half_g * L.h * L.h;
  const float nR = (Fhn - p.half_g * f.hR * f.hR) + p.half_g * R.h * R.h;
  const float tx = Fht * nx, ty = Fht * ny;
  const float w5 = Smax * len;
  lw = make_float4(Fh * len, fmaf(nL, nx, -ty) * len, fmaf(nL, ny, tx) * len, w5);
  rw = make_float4(Fh * -len, fmaf(nR, nx, -ty) * -len, fmaf(nR, ny, tx) * -len, w5);
}

// The edge's CFL wave speed, max(|SL|, |SR|), from the same face states and
// speeds as its words (reference flux.py:172-182, resident.py:528-610).
__device__ __forceinline__ float wave_bound(const Phys& p, const Cell& L, const Cell& R, bool mirror, float d,
                                            float nx, float ny) {
  const Face f = face(p, L, R, mirror, d, nx, ny);
  float SL, SR;
  speeds(p, f, SL, SR);
  return vmax(fabsf(SL), fabsf(SR));
}

// Film drag 1 / (1 + dt (e^2 - h^2) / (2h^2) sqrt(g/e)) below e_film
// (reference flux.py:46-51), for dry_tol <= h and dt >= 0.
__device__ __forceinline__ float film_drag(const Phys& p, float h, float dt) {
  if (h >= p.e_film) return 1.0f;
  const float hh = h * h;
  const float rate = ((p.e_film2 - hh) * rcp(2.0f * hh)) * p.sqrt_g_e;
  return rcp(vmin(fmaf(dt, rate, 1.0f), kRcpMax));
}

// x^(-1/3) for positive normal x: exact exponent split, cubic start on
// [1, 8), three Newton steps y <- y (4/3 - (x/3) y^3); no division.
__device__ __forceinline__ float rcbrt(float x) {
  const uint32_t u = __float_as_uint(x);
  const int e = (int)((u >> 23) & 0xffu) - 127;
  const int k = floor_div3(e);
  const int r = e - 3 * k;
  const float m = __uint_as_float((u & 0x7fffffu) | ((uint32_t)(127 + r) << 23));
  float y = fmaf(fmaf(fmaf(-0.00191212f, m, 0.03603631f), m, -0.25165805f), m, 1.17793f);
  const float m3 = m * 0.33333334f;
  for (int it = 0; it < 3; ++it) y = fmaf(-m3, y * y * y, 1.3333334f) * y;
  return y * __uint_as_float((uint32_t)(127 - k) << 23);
}

// Semi-implicit Manning factor 1 / (1 + dt g n^2 |u| h^(-4/3)) with film drag
// applied first; (hu, hv) updated in place for a wet cell. |u| is
// |(hu, hv)| h^-1 with h^-1 taken as r^3 from the inverse cube root r that
// the h^(-4/3) factor needs anyway: no division.
__device__ __forceinline__ void friction(const Phys& p, float h, float n, float dt, float& hu, float& hv) {
  const float drag = fast::film_drag(p, h, dt);
  hu *= drag;
  hv *= drag;
  const float r = rcbrt(h);
  const float r2 = r * r;
  const float spd = sqrt_rn(fmaf(hu, hu, hv * hv)) * (r2 * r);
  const float Sf = p.g * n * n * spd * (r2 * r2);
  const float id = rcp(vmin(fmaf(dt, Sf, 1.0f), kRcpMax));
  hu *= id;
  hv *= id;
}

// Speed |u| and celerity c of a wetted source cell (h >= dry_tol) for its dt
// bound (reference resident.py:709-728).
__device__ __forceinline__ void src_speeds(const Phys& p, float h, float hu, float hv, float& u, float& c) {
  u = sqrt_rn(fmaf(hu, hu, hv * hv)) * rcp(vmin(h, kRcpMax));
  c = p.sqrt_g * sqrt_rn(h);
}

}  // namespace fast

}  // namespace inu
