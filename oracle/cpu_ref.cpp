// FP32 CPU restatement of the reference's GPU-resident shallow-water step.
//
// TEST INFRASTRUCTURE: the oracle the HIP backend is checked against, and the
// timed CPU baseline of the benchmark. It is never part of the product path.
//
// It follows the structure of the reference kernels literally, kernel by
// kernel, so a bitwise match with the fused HIP kernels also checks that the
// fusion changed nothing:
//   flux_resident_gather     inundation/vulkan/shaders/resident.py:187-287
//                            (workgroup skip over erange :233-239, predictor
//                            rule :209-226, wall ghost :126-129, dry skip
//                            :140-145, per-edge words :165-184)
//   update_resident_gather[_comp]  resident.py:401-516 (sum :312-344, depth
//                            :359-373, store :375-398)
//   source_resident          resident.py:665-733
//   block_mark / dilate      resident.py:738-805
//   step_control             resident.py:813-885
//   cfl_edge_gather / cfl_gather / dt_reset   resident.py:528-610, cfl.py:130-166
//   record order             inundation/vulkan/gpu_resident_batch.py:489-513, 567-599
//   HLLC, reconstruction, film drag, two-sum   inundation/vulkan/shaders/flux.py:43-185, 301-320
//   open boundaries: open_forced / inflow_depth / ghost_state /
//                            boundary_mass_flux flux.py:206-281 and the
//                            outflow-only redo OPEN_FLUX_MASS_GLSL :284-295;
//                            the open ghost and forced-edge skip in the flux
//                            kernel resident.py:126-145; the prologue ghost
//                            :556-566; the stretch tally :613-658; forced
//                            blocks gpu_resident_batch.py:447-458
//
// Two arithmetic orders, as in the backend (flag 16 selects the first): the
// reference's literal order (one IEEE division or sqrt per occurrence, no
// fused multiply-add), and the backend's default fast order (namespace fastm:
// shared reciprocals, explicit fma), algebraically the same scheme.
//
// One statement differs from the reference and is a choice, not a port: the
// source kernel's dt bound reads each neighbour's depth after *every* source
// of the step has been applied (the reference reads it while other source
// threads may still be writing it, so its result there depends on thread
// timing). Friction's h^(4/3) is h * cbrt(h) with the IEEE-only cube root of
// the backend, restated here; so are the open-boundary powers: h^(2/3) is
// cbrt(h)^2, and an inflow's depth is a fixed four-step Newton solve.
//
// Build: see oracle/build.py (g++ -O2 -fopenmp -ffp-contract=off
// -fno-fast-math -fno-math-errno). FP32 arithmetic, no contraction.

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kBlk = 256;
constexpr uint32_t kWg = 256;
constexpr float kTicks = 1048576.0f;
constexpr float kSentinel = 1e30f;
constexpr uint32_t kWords = 6;  // EF_WORDS
constexpr uint32_t kCfl = 5;    // EF_CFL

inline uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
// GLSL min(x, y) = y < x ? y : x; max(x, y) = x < y ? y : x.
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline float gclamp(float x, float a, float b) { return gmin(gmax(x, a), b); }

struct Ctrl {
  uint64_t now_ticks, stop_ticks, steps;
  uint32_t dt_ticks, dt_max_ticks, dry_dt_ticks;
  float last_raw, cfl_safety;
  uint32_t pad;
};
static_assert(sizeof(Ctrl) == 48, "ctrl layout");

struct Pc {
  float g, dry, cfl;
};

// ---- flux.py:63-182 -------------------------------------------------------

inline float safe_vel(const Pc& pc, float h_, float hq) {
  float e_ = 4.0f * pc.dry;  // FILM_DEPTHS * dry_tol
  return (h_ > pc.dry) ? 2.0f * h_ * hq / (h_ * h_ + gmax(h_ * h_, e_ * e_)) : 0.0f;
}

inline void hllc_speeds(const Pc& pc, float hL, float unL, float hR, float unR, float& SL, float& SR) {
  float cL = std::sqrt(pc.g * gmax(hL, 0.0f));
  float cR = std::sqrt(pc.g * gmax(hR, 0.0f));
  float sqL = std::sqrt(gmax(hL, 0.0f));
  float sqR = std::sqrt(gmax(hR, 0.0f));
  float den = sqL + sqR;
  float u_roe = (den > 1e-12f) ? (sqL * unL + sqR * unR) / den : 0.0f;
  float c_roe = std::sqrt(pc.g * 0.5f * (hL + hR));
  SL = gmin(unL - cL, u_roe - c_roe);
  SR = gmax(unR + cR, u_roe + c_roe);
  float wetL = gclamp(hL / pc.dry - 1.0f, 0.0f, 1.0f);
  float wetR = gclamp(hR / pc.dry - 1.0f, 0.0f, 1.0f);
  float dryLeft = (1.0f - wetL) * wetR;
  float dryRight = (1.0f - wetR) * wetL;
  float roe = 1.0f - dryLeft - dryRight;
  SL = roe * SL + dryLeft * (unR - 2.0f * cR) + dryRight * (unL - cL);
  SR = roe * SR + dryLeft * (unR + cR) + dryRight * (unL + 2.0f * cL);
}

inline void hllc(const Pc& pc, float hL, float hqnL, float hqtL, float hR, float hqnR, float hqtR,
                 float& Fh, float& Fhn, float& Fht, float& Smax) {
  float unL = safe_vel(pc, hL, hqnL);
  float utL = safe_vel(pc, hL, hqtL);
  float unR = safe_vel(pc, hR, hqnR);
  float utR = safe_vel(pc, hR, hqtR);
  float SL, SR;
  hllc_speeds(pc, hL, unL, hR, unR, SL, SR);
  Smax = gmax(std::fabs(SL), std::fabs(SR));
  float num_sm = SL * hR * (unR - SR) - SR * hL * (unL - SL);
  float den_sm = hR * (unR - SR) - hL * (unL - SL);
  float SM = (std::fabs(den_sm) > 1e-12f) ? num_sm / den_sm : 0.0f;
  float FL_h = hL * unL;
  float FL_hn = hL * unL * unL + 0.5f * pc.g * hL * hL;
  float FL_ht = hL * unL * utL;
  float FR_h = hR * unR;
  float FR_hn = hR * unR * unR + 0.5f * pc.g * hR * hR;
  float FR_ht = hR * unR * utR;
  float dSL = SL - SM;
  float dSR = SR - SM;
  float kL = hL * (SL - unL) / ((std::fabs(dSL) > 1e-14f) ? dSL : 1e-14f);
  float kR = hR * (SR - unR) / ((std::fabs(dSR) > 1e-14f) ? dSR : 1e-14f);
  float FsL_h = FL_h + SL * (kL - hL);
  float FsL_hn = FL_hn + SL * (kL * SM - hL * unL);
  float FsL_ht = FL_ht + SL * (kL * utL - hL * utL);
  float FsR_h = FR_h + SR * (kR - hR);
  float FsR_hn = FR_hn + SR * (kR * SM - hR * unR);
  float FsR_ht = FR_ht + SR * (kR * utR - hR * utR);
  if (SL >= 0.0f) { Fh = FL_h; Fhn = FL_hn; Fht = FL_ht; }
  else if (SM >= 0.0f) { Fh = FsL_h; Fhn = FsL_hn; Fht = FsL_ht; }
  else if (SR >= 0.0f) { Fh = FsR_h; Fhn = FsR_hn; Fht = FsR_ht; }
  else { Fh = FR_h; Fhn = FR_hn; Fht = FR_ht; }
}

inline void face_states(const Pc& pc, float hL, float huL, float hvL, float dzL, float hR, float huR,
                        float hvR, float dzR, float nx_, float ny_, float& hLs, float& hqnL,
                        float& hqtL, float& hRs, float& hqnR, float& hqtR) {
  hLs = gmax(0.0f, hL - dzL);
  hRs = gmax(0.0f, hR - dzR);
  float uL = safe_vel(pc, hL, huL);
  float vL = safe_vel(pc, hL, hvL);
  float uR = safe_vel(pc, hR, huR);
  float vR = safe_vel(pc, hR, hvR);
  float huLs = hLs * uL;
  float hvLs = hLs * vL;
  float huRs = hRs * uR;
  float hvRs = hRs * vR;
  hqnL = huLs * nx_ + hvLs * ny_;
  hqtL = -huLs * ny_ + hvLs * nx_;
  hqnR = huRs * nx_ + hvRs * ny_;
  hqtR = -huRs * ny_ + hvRs * nx_;
}

inline float wave_bound(const Pc& pc, float hL, float huL, float hvL, float dzL, float hR, float huR,
                        float hvR, float dzR, float nx_, float ny_) {
  float hLs, hqnL, hqtL, hRs, hqnR, hqtR;
  face_states(pc, hL, huL, hvL, dzL, hR, huR, hvR, dzR, nx_, ny_, hLs, hqnL, hqtL, hRs, hqnR, hqtR);
  float SL, SR;
  hllc_speeds(pc, hLs, safe_vel(pc, hLs, hqnL), hRs, safe_vel(pc, hRs, hqnR), SL, SR);
  return gmax(std::fabs(SL), std::fabs(SR));
}

inline float film_drag(const Pc& pc, float h_, float dt_) {
  float e_ = 4.0f * pc.dry;
  if (h_ >= e_) return 1.0f;
  float rate = (e_ * e_ - h_ * h_) / (2.0f * h_ * h_) * std::sqrt(pc.g / e_);
  return 1.0f / (1.0f + dt_ * rate);
}

// ---- flux.py:301-319 -------------------------------------------------------

inline void two_sum(float a, float b, float& s, float& e) {
  float t = a + b;
  float bb = t - a;
  float err = (a - (t - bb)) + (b - bb);
  s = t;
  e = err;
}

inline void add_compensated(float hi, float loi, float inc, float& h_new, float& lo_new) {
  float s, e;
  two_sum(hi, inc, s, e);
  e += loi;
  two_sum(s, e, h_new, lo_new);
  if (h_new < 0.0f || (h_new == 0.0f && lo_new < 0.0f)) {
    h_new = 0.0f;
    lo_new = 0.0f;
  }
}

// h^(4/3) = h * cbrt(h): exact exponent split, cubic start on [1, 8), two
// Halley steps (the backend's IEEE-only cube root, restated).
inline float cbrt_ieee(float x) {
  uint32_t u = fbits(x);
  int e = (int)((u >> 23) & 0xffu) - 127;
  int k = (e >= 0) ? e / 3 : -((2 - e) / 3);
  int r = e - 3 * k;
  float m = bitsf((u & 0x7fffffu) | ((uint32_t)(127 + r) << 23));
  float y = 0.80169147f + m * (0.24778609f - m * 0.0127262f);
  float t = y * y * y;
  y = y * (t + 2.0f * m) / (2.0f * t + m);
  t = y * y * y;
  y = y * (t + 2.0f * m) / (2.0f * t + m);
  return y * bitsf((uint32_t)(127 + k) << 23);
}
inline float pow43(float h) { return h * cbrt_ieee(h); }

// ---- open boundaries (flux.py:188-272) ------------------------------------

constexpr uint32_t kOpenWall = 0u, kOpenFree = 1u, kOpenNormal = 2u, kOpenStage = 3u, kOpenInflow = 4u;

// h^(2/3) as cbrt(h)^2; 0 below the smallest normal float.
inline float pow23(float h) {
  if (!(h >= 1.17549435e-38f)) return 0.0f;
  float c = cbrt_ieee(h);
  return c * c;
}

// Depth beyond an inflow edge of discharge q > 0 per unit width: the root of
// -q/hG + 2 sqrt(g hG) = riemann (the outgoing characteristic), Newton in
// s = sqrt(hG) from the larger of the critical-flow and still-water starts,
// four steps, each keeping at least half of s (flux.py:209-223), the cube root
// being the IEEE-only one above.
inline float inflow_depth(const Pc& pc, float q, float riemann) {
  const float root_g = std::sqrt(pc.g);
  float s = gmax(cbrt_ieee(q / (2.0f * root_g)), riemann / (2.0f * root_g));
  for (int k = 0; k < 4; ++k) {
    float f = 2.0f * root_g * s - q / (s * s) - riemann;
    float df = 2.0f * root_g + 2.0f * q / (s * s * s);
    s = gmax(s - f / df, 0.5f * s);
  }
  return s * s;
}

// flux.py:224-256: the state beyond an open edge; value is sqrt(S)/n for
// normal depth, the depth over the cell's bed for a stage, the discharge per
// unit width for an inflow.
inline void ghost_state(const Pc& pc, uint32_t kind, float value, float hL, float huL, float hvL, float nx,
                        float ny, float& hG, float& huG, float& hvG) {
  float uL = safe_vel(pc, hL, huL);
  float vL = safe_vel(pc, hL, hvL);
  float unL = uL * nx + vL * ny;
  float ut = -uL * ny + vL * nx;
  float cL = std::sqrt(pc.g * gmax(hL, 0.0f));
  float un = -unL;
  hG = hL;
  if (kind == kOpenFree) {
    un = std::fabs(unL);
  } else if (kind == kOpenNormal) {
    un = value * pow23(gmax(hL, 0.0f));
  } else if (kind == kOpenStage) {
    if (hL < pc.dry || unL < cL) {
      hG = value;
      un = unL + 2.0f * (cL - std::sqrt(pc.g * value));
    } else {
      un = unL;
    }
  } else if (kind == kOpenInflow && value > 0.0f) {
    hG = inflow_depth(pc, value, unL + 2.0f * cL);
    un = -value / hG;
    ut = 0.0f;
  }
  huG = hG * (un * nx - ut * ny);
  hvG = hG * (un * ny + ut * nx);
}

inline bool open_forced(uint32_t kind, float value) { return kind == kOpenInflow && value > 0.0f; }
inline bool outflow_only(uint32_t kind) { return kind == kOpenFree || kind == kOpenNormal; }

// flux.py:271-281: mass flux through an open edge of HLLC mass flux fh.
inline float boundary_mass_flux(uint32_t kind, float value, float fh) {
  if (kind == kOpenFree || kind == kOpenNormal) return gmax(fh, 0.0f);
  if (kind == kOpenStage) return fh;
  if (kind == kOpenInflow) return -value;
  return 0.0f;
}

// ---- the backend's fast arithmetic, restated -------------------------------
// Same scheme, algebraically unchanged (inundation_amd/csrc/swe_math.h,
// namespace fast):
//  - per cell the desingularised velocity (u, v) = (hu, hv) vf(h) and the
//    face factor phi(h) = h vf(h), with vf(h) = 1/h for h >= e_film (so phi
//    is exactly 1 there) and 2h / (h^2 + e^2) below; a side's face velocity
//    is its cell velocity in the edge frame times phi(face depth);
//  - a wall's mirror taken in the edge frame (normal velocity negated);
//  - divisions as products with the IEEE reciprocal 1/x of a divisor kept in
//    [2^-126, 2^126] (the backend forms it from the hardware estimate and one
//    Newton step, equal to 1/x over that range);
//  - min, max and the [0, 1] clamp with the hardware's semantics;
//  - c = sqrt(g) sqrt(h), h * (1 / dry_tol), one star state chosen before it
//    is formed, explicit fused multiply-adds, h^(-4/3) from a division-free
//    inverse cube root.
namespace fastm {

struct K {
  float g, dry, cfl, e, e2, half_g, sqrt_g_e, inv_dry, sqrt_g;
};

inline K consts(const Pc& pc) {
  K k;
  k.g = pc.g;
  k.dry = pc.dry;
  k.cfl = pc.cfl;
  k.e = 4.0f * pc.dry;
  k.e2 = k.e * k.e;
  k.half_g = 0.5f * pc.g;
  k.sqrt_g_e = std::sqrt(pc.g / k.e);
  k.inv_dry = 1.0f / pc.dry;
  k.sqrt_g = std::sqrt(pc.g);
  return k;
}

constexpr float kRcpMax = 0x1p126f;

inline float rcp(float x) { return 1.0f / x; }

// The hardware's v_min_f32 / v_max_f32: a NaN operand yields the other, and
// -0 orders below +0.
inline float vmin(float x, float y) {
  if (x != x) return y;
  if (y != y) return x;
  if (x == 0.0f && y == 0.0f) return std::signbit(x) ? x : y;
  return (y < x) ? y : x;
}
inline float vmax(float x, float y) {
  if (x != x) return y;
  if (y != y) return x;
  if (x == 0.0f && y == 0.0f) return std::signbit(x) ? y : x;
  return (x < y) ? y : x;
}
// The [0, 1] clamp (v_med3_f32 / the clamp modifier): NaN and -0 give +0.
inline float clamp01(float x) { return (x > 0.0f) ? ((x < 1.0f) ? x : 1.0f) : 0.0f; }
// v_med3_f32(x, -2^126, 2^126): NaN gives -2^126.
inline float rcp_range(float x) {
  if (x != x) return -kRcpMax;
  return (x < -kRcpMax) ? -kRcpMax : ((x > kRcpMax) ? kRcpMax : x);
}

inline float thin_phi(const K& k, float h) {
  float hh = h * h;
  return (h > k.dry) ? (2.0f * hh) * rcp(hh + k.e2) : 0.0f;
}

inline void factors(const K& k, float h, float& vf, float& phi) {
  if (!(h > k.dry)) {
    vf = 0.0f;
    phi = 0.0f;
  } else if (h >= k.e) {
    vf = rcp(vmin(h, kRcpMax));
    phi = 1.0f;
  } else {
    float hh = h * h;
    float r = rcp(hh + k.e2);
    vf = (2.0f * h) * r;
    phi = (2.0f * hh) * r;
  }
}

struct Cell {
  float h, u, v, lo, phi, sq;
};

inline Cell cell(const K& k, float h, float hu, float hv, float lo) {
  float vf, phi;
  factors(k, h, vf, phi);
  return Cell{h, hu * vf, hv * vf, lo, phi, std::sqrt(vmax(h, 0.0f))};
}

struct Face {
  float hL, hR, sqL, sqR, unL, utL, unR, utR;
};

// d > 0: the left face steps up by d; d < 0: the right one by -d.
inline Face face(const K& k, const Cell& L, const Cell& R, bool mirror, float d, float nx, float ny) {
  Face f;
  float t = ((L.h - R.h) - d) + (L.lo - R.lo);
  bool stepL = d > 0.0f, stepR = d < 0.0f;
  f.hL = stepL ? vmax(R.h + t, 0.0f) : L.h;
  f.hR = stepR ? vmax(L.h - t, 0.0f) : R.h;
  float hs = stepL ? f.hL : f.hR;
  float sqs = std::sqrt(hs);
  float phs = (hs >= k.e) ? 1.0f : thin_phi(k, hs);
  f.sqL = stepL ? sqs : L.sq;
  f.sqR = stepR ? sqs : R.sq;
  float phL = stepL ? phs : L.phi, phR = stepR ? phs : R.phi;
  float nL = std::fma(L.u, nx, L.v * ny);
  float tL = std::fma(L.v, nx, -(L.u * ny));
  float nR = mirror ? -nL : std::fma(R.u, nx, R.v * ny);
  float tR = mirror ? tL : std::fma(R.v, nx, -(R.u * ny));
  f.unL = nL * phL;
  f.utL = tL * phL;
  f.unR = nR * phR;
  f.utR = tR * phR;
  return f;
}

inline void speeds(const K& k, const Face& f, float& SL, float& SR) {
  float cL = k.sqrt_g * f.sqL, cR = k.sqrt_g * f.sqR;
  float den = f.sqL + f.sqR;
  float u_roe = (den > 1e-12f) ? std::fma(f.sqL, f.unL, f.sqR * f.unR) * rcp(vmin(den, kRcpMax)) : 0.0f;
  float c_roe = std::sqrt(k.half_g * (f.hL + f.hR));
  float sl = vmin(f.unL - cL, u_roe - c_roe);
  float sr = vmax(f.unR + cR, u_roe + c_roe);
  float wetL = clamp01(std::fma(f.hL, k.inv_dry, -1.0f));
  float wetR = clamp01(std::fma(f.hR, k.inv_dry, -1.0f));
  float dryLeft = (1.0f - wetL) * wetR;
  float dryRight = (1.0f - wetR) * wetL;
  float roe = (1.0f - dryLeft) - dryRight;
  SL = std::fma(roe, sl, std::fma(dryLeft, f.unR - 2.0f * cR, dryRight * (f.unL - cL)));
  SR = std::fma(roe, sr, std::fma(dryLeft, f.unR + cR, dryRight * (f.unL + 2.0f * cL)));
}

inline void hllc(const K& k, const Face& f, float& Fh, float& Fhn, float& Fht, float& Smax) {
  float SL, SR;
  speeds(k, f, SL, SR);
  Smax = vmax(std::fabs(SL), std::fabs(SR));
  const float hL = f.hL, hR = f.hR, unL = f.unL, unR = f.unR, utL = f.utL, utR = f.utR;
  // Toro's contact speed with the numerator regrouped as
  // SL SR (hL - hR) + (SL hR unR - SR hL unL): exactly zero for still water.
  float num_sm = std::fma(SL * SR, hL - hR, std::fma(SL * hR, unR, -(SR * hL * unL)));
  float den_sm = std::fma(hR, unR - SR, -(hL * (unL - SL)));
  float SM = (std::fabs(den_sm) > 1e-12f) ? num_sm * rcp(rcp_range(den_sm)) : 0.0f;
  float FL_h = hL * unL, FL_hn = std::fma(FL_h, unL, k.half_g * hL * hL), FL_ht = FL_h * utL;
  float FR_h = hR * unR, FR_hn = std::fma(FR_h, unR, k.half_g * hR * hR), FR_ht = FR_h * utR;
  bool left = SM >= 0.0f;
  float S = left ? SL : SR, hS = left ? hL : hR, unS = left ? unL : unR, utS = left ? utL : utR;
  float FS_h = left ? FL_h : FR_h, FS_hn = left ? FL_hn : FR_hn, FS_ht = left ? FL_ht : FR_ht;
  // Star depth minus the side depth, hS (SM - unS) / (S - SM): zero when the
  // contact does not move relative to the side.
  float dS = S - SM;
  bool dS_ok = std::fabs(dS) > 1e-14f;
  float ks_q = (hS * (dS_ok ? SM - unS : S - unS)) * rcp(dS_ok ? rcp_range(dS) : 1e-14f);
  float ksh = dS_ok ? ks_q : ks_q - hS;
  float kS = hS + ksh;
  if (SL >= 0.0f) { Fh = FL_h; Fhn = FL_hn; Fht = FL_ht; }
  else if (!left && !(SR >= 0.0f)) { Fh = FR_h; Fhn = FR_hn; Fht = FR_ht; }
  else {
    Fh = std::fma(S, ksh, FS_h);
    Fhn = std::fma(S, std::fma(kS, SM, -(hS * unS)), FS_hn);
    Fht = std::fma(S, ksh * utS, FS_ht);
  }
}

// The six words of an edge; `wall` passes no mass, `mirror`: R is L's wall
// mirror.
inline void words(const K& k, bool wall, const Cell& L, const Cell& R, bool mirror, float d, float nx, float ny,
                  float len, float* o, float& fh) {
  Face f = face(k, L, R, mirror, d, nx, ny);
  float Fh, Fhn, Fht, Smax;
  hllc(k, f, Fh, Fhn, Fht, Smax);
  fh = Fh;
  if (wall) Fh = 0.0f;
  // Normal momentum flux plus the hydrostatic correction, grouped as
  // (Fhn - P(face depth)) + P(cell depth): still water gives exactly the
  // cell's own pressure on every edge.
  float nL = (Fhn - k.half_g * f.hL * f.hL) + k.half_g * L.h * L.h;
  float nR = (Fhn - k.half_g * f.hR * f.hR) + k.half_g * R.h * R.h;
  float tx = Fht * nx, ty = Fht * ny;
  o[0] = Fh * len;
  o[1] = std::fma(nL, nx, -ty) * len;
  o[2] = std::fma(nL, ny, tx) * len;
  o[3] = std::fma(nR, nx, -ty) * len;
  o[4] = std::fma(nR, ny, tx) * len;
  o[5] = Smax * len;
}

inline float wave_bound(const K& k, const Cell& L, const Cell& R, bool mirror, float d, float nx, float ny) {
  Face f = face(k, L, R, mirror, d, nx, ny);
  float SL, SR;
  speeds(k, f, SL, SR);
  return vmax(std::fabs(SL), std::fabs(SR));
}

inline float film(const K& k, float h, float dt) {
  if (h >= k.e) return 1.0f;
  float hh = h * h;
  float rate = ((k.e2 - hh) * rcp(2.0f * hh)) * k.sqrt_g_e;
  return rcp(vmin(std::fma(dt, rate, 1.0f), kRcpMax));
}

inline float rcbrt(float x) {
  uint32_t u = fbits(x);
  int e = (int)((u >> 23) & 0xffu) - 127;
  int kk = (e >= 0) ? e / 3 : -((2 - e) / 3);
  int r = e - 3 * kk;
  float m = bitsf((u & 0x7fffffu) | ((uint32_t)(127 + r) << 23));
  float y = std::fma(std::fma(std::fma(-0.00191212f, m, 0.03603631f), m, -0.25165805f), m, 1.17793f);
  float m3 = m * 0.33333334f;
  for (int it = 0; it < 3; ++it) y = std::fma(-m3, y * y * y, 1.3333334f) * y;
  return y * bitsf((uint32_t)(127 - kk) << 23);
}

inline void friction(const K& k, float h, float n, float dt, float& hu, float& hv) {
  float drag = film(k, h, dt);
  hu *= drag;
  hv *= drag;
  float r = rcbrt(h);
  float r2 = r * r;
  float spd = std::sqrt(std::fma(hu, hu, hv * hv)) * (r2 * r);  // |u| with 1/h as r^3
  float Sf = k.g * n * n * spd * (r2 * r2);
  float id = rcp(vmin(std::fma(dt, Sf, 1.0f), kRcpMax));
  hu *= id;
  hv *= id;
}

// |u| and c of a wetted source cell (h >= dry_tol) for its dt bound.
inline void src_speeds(const K& k, float h, float hu, float hv, float& u, float& c) {
  u = std::sqrt(std::fma(hu, hu, hv * hv)) * rcp(vmin(h, kRcpMax));
  c = k.sqrt_g * std::sqrt(h);
}

}  // namespace fastm

}  // namespace

struct cref {
  uint32_t N = 0, E = 0, NB = 0, NWG = 0;
  std::vector<int32_t> ecL, ecR;
  std::vector<float> enx, eny, elen, edz, area;
  std::vector<float> ia;  // 1 / area, for the fast order
  std::vector<uint32_t> cptr, cidx, erange, bptr, bidx;
  // state
  std::vector<float> h, hu, hv, h1, hu1, hv1, hlo, h1lo, nman, clamp, ef;
  std::vector<uint8_t> touched;
  std::vector<uint32_t> blk;  // 3 words per block: wet, active, forced
  std::vector<uint32_t> src_cell;
  std::vector<float> src_rate;
  // open edges: edge e with ecR[e] = -2 - k is open edge k
  std::vector<uint32_t> okind, sptr;
  std::vector<float> oval, tally;  // tally: per stretch out (hi, lo), in (hi, lo)
  std::vector<uint32_t> oedge;     // the edge of open edge k
  uint32_t dtbuf[4] = {0, 0, 0, 0};
  Ctrl ctrl{};
  Pc pc{9.81f, 1e-4f, 0.8f};
  bool compensated = true, active_set = true, track_clamp = false, fast = true;
};

namespace {

inline bool block_active(const cref& r, uint32_t b) { return r.blk[3 * b + 1] != 0u; }

// dt bound bits of cell i over its summed wave speeds lam: cfl A / lam for
// lam > 1e-30 and A > 1e-10 (resident.py:509, :604, :726), else none. The
// fast order takes cfl / (A^-1 lam) with the cell's inverse area, guarded by
// A^-1 < 1e10.
inline uint32_t cell_bound(const cref& r, const Pc& pc, uint32_t i, float lam) {
  if (!(lam > 1e-30f)) return 0xffffffffu;
  if (r.fast) return r.ia[i] < 1e10f ? fbits(pc.cfl / (r.ia[i] * lam)) : 0xffffffffu;
  return r.area[i] > 1e-10f ? fbits(pc.cfl * r.area[i] / lam) : 0xffffffffu;
}

// Low word of the stage input depth (compensated runs; zero otherwise).
inline float load_lo(const cref& r, int stage, int c) {
  if (!r.compensated) return 0.0f;
  return (stage == 1 && block_active(r, (uint32_t)c >> 8)) ? r.h1lo[c] : r.hlo[c];
}

inline void load_state(const cref& r, int stage, int c, float& h_, float& hu_, float& hv_) {
  if (stage == 1 && block_active(r, (uint32_t)c >> 8)) {
    h_ = r.h1[c]; hu_ = r.hu1[c]; hv_ = r.hv1[c];
  } else {
    h_ = r.h[c]; hu_ = r.hu[c]; hv_ = r.hv[c];
  }
}

// The six words of one edge (resident.py:146-184): mass, left momentum x, y,
// right momentum x, y, CFL term; fh receives the HLLC mass flux, and a wall
// passes no mass. `mirror`: the right side is the left cell's wall mirror
// (the exact order gets its mirrored momenta in huR, hvR; the fast order
// takes the mirror in the edge frame).
void edge_words(const cref& r, const fastm::K& k, bool wall, bool mirror, float hL, float huL, float hvL, float loL,
                float hR, float huR, float hvR, float loR, float dzL, float dzR, float nx_, float ny_, float len,
                float* o, float& fh) {
  const Pc& pc = r.pc;
  if (r.fast) {
    const fastm::Cell L = fastm::cell(k, hL, huL, hvL, loL);
    const fastm::Cell R = mirror ? L : fastm::cell(k, hR, huR, hvR, loR);
    const float d = (dzL > 0.0f) ? dzL : -dzR;  // the backend's signed step
    fastm::words(k, wall, L, R, mirror, d, nx_, ny_, len, o, fh);
    return;
  }
  float hLs, hqnL, hqtL, hRs, hqnR, hqtR;
  face_states(pc, hL, huL, hvL, dzL, hR, huR, hvR, dzR, nx_, ny_, hLs, hqnL, hqtL, hRs, hqnR, hqtR);
  float Fh, Fhn, Fht, Smax;
  hllc(pc, hLs, hqnL, hqtL, hRs, hqnR, hqtR, Fh, Fhn, Fht, Smax);
  fh = Fh;
  if (wall) Fh = 0.0f;  // boundary_mass_flux of a wall
  float Fhu = Fhn * nx_ - Fht * ny_;
  float Fhv = Fhn * ny_ + Fht * nx_;
  float SL_g = 0.5f * pc.g * (hL * hL - hLs * hLs);
  float SR_g = 0.5f * pc.g * (hR * hR - hRs * hRs);
  o[0] = Fh * len;
  o[1] = Fhu * len + SL_g * nx_ * len;
  o[2] = Fhv * len + SL_g * ny_ * len;
  o[3] = Fhu * len + SR_g * nx_ * len;
  o[4] = Fhv * len + SR_g * ny_ * len;
  o[5] = Smax * len;
}

void flux(cref& r, int stage) {
  if (r.dtbuf[0] == 0u) return;
  const Pc pc = r.pc;
  const fastm::K kk = fastm::consts(pc);
#pragma omp parallel for schedule(static)
  for (int64_t wg = 0; wg < (int64_t)r.NWG; ++wg) {
    bool live = false;
    for (uint32_t b = r.erange[2 * wg]; b <= r.erange[2 * wg + 1]; ++b)
      if (block_active(r, b)) { live = true; break; }
    if (!live) continue;
    const uint32_t e1 = std::min<uint32_t>((uint32_t)(wg + 1) * kWg, r.E);
    for (uint32_t eid = (uint32_t)wg * kWg; eid < e1; ++eid) {
      int cL = r.ecL[eid], cR = r.ecR[eid];
      float nx_ = r.enx[eid], ny_ = r.eny[eid], len = r.elen[eid];
      float hL, huL, hvL, hR, huR, hvR;
      load_state(r, stage, cL, hL, huL, hvL);
      const float loL = load_lo(r, stage, cL);
      const float loR = (cR >= 0) ? load_lo(r, stage, cR) : loL;
      const bool open = cR < -1;
      const uint32_t ok = open ? (uint32_t)(-2 - cR) : 0u;
      if (cR == -1) {
        hR = hL;
        huR = huL - 2.0f * (huL * nx_ + hvL * ny_) * nx_;
        hvR = hvL - 2.0f * (huL * nx_ + hvL * ny_) * ny_;
      } else if (open) {
        ghost_state(pc, r.okind[ok], r.oval[ok], hL, huL, hvL, nx_, ny_, hR, huR, hvR);
      } else {
        load_state(r, stage, cR, hR, huR, hvR);
      }
      float* o = &r.ef[(size_t)kWords * eid];
      if (hL < pc.dry && hR < pc.dry && !(open && open_forced(r.okind[ok], r.oval[ok]))) {
        // An open edge clears its words; the gather skips other dry edges.
        if (open)
          for (uint32_t w = 0; w < kWords; ++w) o[w] = 0.0f;
        continue;
      }
      float dzL = r.edz[2 * eid], dzR = r.edz[2 * eid + 1];
      float w[kWords], fh;
      edge_words(r, kk, cR == -1, cR == -1, hL, huL, hvL, loL, hR, huR, hvR, loR, dzL, dzR, nx_, ny_, len, w, fh);
      if (open) {
        const uint32_t kind = r.okind[ok];
        if (outflow_only(kind) && fh < 0.0f) {
          // Drawing water in through an outflow-only edge: the whole flux is
          // redone against the wall ghost and passes no mass.
          ghost_state(pc, kOpenWall, 0.0f, hL, huL, hvL, nx_, ny_, hR, huR, hvR);
          edge_words(r, kk, true, false, hL, huL, hvL, loL, hR, huR, hvR, loR, dzL, dzR, nx_, ny_, len, w, fh);
        } else {
          w[0] = boundary_mass_flux(kind, r.oval[ok], fh) * len;
        }
      }
      __atomic_store_n(&r.touched[cL], (uint8_t)1, __ATOMIC_RELAXED);
      if (cR >= 0) __atomic_store_n(&r.touched[cR], (uint8_t)1, __ATOMIC_RELAXED);
      for (uint32_t q = 0; q < 5; ++q) o[q] = w[q];
      if (stage == 1) o[kCfl] = w[5];
    }
  }
}

// Volume through each stretch after one flux stage (resident.py:613-658):
// dt/2 of each open edge's mass word, out and in apart, each kept as an exact
// pair of words; an edge of an inactive block moved nothing.
void open_tally(cref& r) {
  if (r.dtbuf[0] == 0u || r.oedge.empty()) return;
  const float half_dt = 0.5f * bitsf(r.dtbuf[0]);
  const uint32_t n_st = (uint32_t)r.sptr.size() - 1;
  for (uint32_t st = 0; st < n_st; ++st) {
    for (uint32_t k = r.sptr[st]; k < r.sptr[st + 1]; ++k) {
      const uint32_t e = r.oedge[k];
      if (!block_active(r, (uint32_t)r.ecL[e] >> 8)) continue;
      float v = half_dt * r.ef[(size_t)kWords * e];
      uint32_t o = 4u * st + ((v < 0.0f) ? 2u : 0u);
      float sum, err, hi, lo;
      two_sum(r.tally[o], std::fabs(v), sum, err);
      err += r.tally[o + 1];
      two_sum(sum, err, hi, lo);
      r.tally[o] = hi;
      r.tally[o + 1] = lo;
    }
  }
}

void update(cref& r, int stage) {
  if (r.dtbuf[0] == 0u) return;
  const Pc pc = r.pc;
  const fastm::K kk = fastm::consts(pc);
  const float dt = bitsf(r.dtbuf[0]);
  if (stage == 0)
    for (uint32_t b = 0; b < r.NB; ++b)
      if (block_active(r, b)) r.blk[3 * b + 0] = 0u;
  uint32_t best = 0xffffffffu;
#pragma omp parallel for schedule(static) reduction(min : best)
  for (int64_t ii = 0; ii < (int64_t)r.N; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const uint32_t b = i >> 8;
    if (!block_active(r, b)) continue;
    float A = r.area[i];
    // The fast order multiplies by the cell's inverse area.
    float dti = r.fast ? dt * r.ia[i] : dt / A;
    float hi, hui, hvi;
    if (stage == 0) { hi = r.h[i]; hui = r.hu[i]; hvi = r.hv[i]; }
    else { hi = r.h1[i]; hui = r.hu1[i]; hvi = r.hv1[i]; }
    // _UPDATE_SUM_GATHER
    float sdh = 0.0f, sdhu = 0.0f, sdhv = 0.0f, lam = 0.0f;
    uint32_t kend = r.cptr[i];
    if (r.touched[i] != 0u) {
      r.touched[i] = 0u;
      kend = r.cptr[i + 1];
    }
    bool own_dry = hi < pc.dry;
    for (uint32_t k = r.cptr[i]; k < kend; ++k) {
      uint32_t s = r.cidx[k];
      uint32_t e = s >> 1;
      bool right = (s & 1u) != 0u;
      if (own_dry) {
        int nb = right ? r.ecL[e] : r.ecR[e];
        if (nb == -1) continue;
        if (nb >= 0) {
          bool pred = stage == 1 && block_active(r, (uint32_t)nb >> 8);
          if ((pred ? r.h1[nb] : r.h[nb]) < pc.dry) continue;
        }
      }
      const float* o = &r.ef[(size_t)kWords * e];
      if (right) { sdh -= o[0]; sdhu -= o[3]; sdhv -= o[4]; }
      else { sdh += o[0]; sdhu += o[1]; sdhv += o[2]; }
      if (stage == 1) lam += o[kCfl];
    }
    float h_new, lo_new, h_raw;
    if (r.compensated) {
      float loi = (stage == 0) ? r.hlo[i] : r.h1lo[i];
      add_compensated(hi, loi, -dti * sdh, h_new, lo_new);
      float s_raw, e_raw;
      two_sum(hi, -dti * sdh, s_raw, e_raw);
      h_raw = s_raw + (e_raw + loi);
    } else {
      h_raw = r.fast ? std::fma(-dti, sdh, hi) : hi - dti * sdh;
      h_new = gmax(h_raw, 0.0f);
      lo_new = 0.0f;
    }
    float hu_new = r.fast ? std::fma(-dti, sdhu, hui) : hui - dti * sdhu;
    float hv_new = r.fast ? std::fma(-dti, sdhv, hvi) : hvi - dti * sdhv;
    if (r.track_clamp) r.clamp[i] += 0.5f * A * (h_new - h_raw);
    if (h_new < pc.dry) { hu_new = 0.0f; hv_new = 0.0f; }
    if (r.fast && h_new >= pc.dry) {
      fastm::friction(kk, h_new, r.nman[i], dt, hu_new, hv_new);
    } else if (h_new >= pc.dry) {
      float drag = film_drag(pc, h_new, dt);
      hu_new *= drag;
      hv_new *= drag;
    }
    if (!r.fast && h_new >= pc.dry) {
      float n = r.nman[i];
      float u_ = hu_new / h_new;
      float v_ = hv_new / h_new;
      float spd = std::sqrt(u_ * u_ + v_ * v_);
      float h43 = pow43(h_new);
      float Sf = pc.g * n * n * spd / h43;
      float den = 1.0f + dt * Sf;
      hu_new = hu_new / den;
      hv_new = hv_new / den;
    }
    float hf = 0.0f;
    if (r.compensated) {
      if (stage == 0) {
        r.h1[i] = h_new;
        r.h1lo[i] = lo_new;
      } else {
        float s, e;
        two_sum(r.h[i], h_new, s, e);
        e += r.hlo[i] + lo_new;
        float lof;
        two_sum(0.5f * s, 0.5f * e, hf, lof);
        r.h[i] = hf;
        r.hlo[i] = lof;
      }
    } else {
      if (stage == 0) r.h1[i] = h_new;
      else {
        hf = 0.5f * (r.h[i] + h_new);
        r.h[i] = hf;
      }
    }
    if (stage == 0) {
      r.hu1[i] = hu_new;
      r.hv1[i] = hv_new;
    } else {
      float huf = 0.5f * (r.hu[i] + hu_new);
      float hvf = 0.5f * (r.hv[i] + hv_new);
      if (hf < pc.dry) { huf = 0.0f; hvf = 0.0f; }
      r.hu[i] = huf;
      r.hv[i] = hvf;
      if (hf >= pc.dry) __atomic_store_n(&r.blk[3 * b + 0], 1u, __ATOMIC_RELAXED);
      best = std::min(best, cell_bound(r, pc, i, lam));
    }
  }
  if (stage == 1) r.dtbuf[1] = std::min(r.dtbuf[1], best);
}

void source(cref& r) {
  if (r.dtbuf[0] == 0u) return;
  const Pc pc = r.pc;
  const float dt = bitsf(r.dtbuf[0]);
  const int64_t n = (int64_t)r.src_cell.size();
  // Every source first ...
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    uint32_t i = r.src_cell[k];
    float inc = dt * r.src_rate[k];
    float h_new;
    if (r.compensated) {
      float lo_new;
      add_compensated(r.h[i], r.hlo[i], inc, h_new, lo_new);
      r.hlo[i] = lo_new;
    } else {
      h_new = gmax(r.h[i] + inc, 0.0f);
    }
    r.h[i] = h_new;
    if (h_new >= pc.dry) __atomic_store_n(&r.blk[3 * (i >> 8) + 0], 1u, __ATOMIC_RELAXED);
  }
  // ... then the post-source dt bound of each wet source cell. The fast order
  // takes (u + c) times the cell's perimeter (its edge lengths summed in CSR
  // order) where the cell's 256-cell block is wet throughout after the
  // sources and no neighbour is dry: the same sum, distributed.
  std::vector<uint8_t> all_wet(r.NB, 1u);
  for (uint32_t i = 0; i < r.N; ++i)
    if (!(r.h[i] >= pc.dry)) all_wet[i >> 8] = 0u;
  uint32_t best = 0xffffffffu;
#pragma omp parallel for schedule(static) reduction(min : best)
  for (int64_t k = 0; k < n; ++k) {
    uint32_t i = r.src_cell[k];
    float h_new = r.h[i];
    if (!(h_new >= pc.dry)) continue;
    const bool F = r.fast;
    float u, c;
    if (F) {
      fastm::src_speeds(fastm::consts(pc), h_new, r.hu[i], r.hv[i], u, c);
    } else {
      u = std::sqrt(r.hu[i] * r.hu[i] + r.hv[i] * r.hv[i]) / h_new;
      c = std::sqrt(pc.g * h_new);
    }
    float lam = 0.0f, perim = 0.0f;
    bool any_dry = false;
    for (uint32_t kk = r.cptr[i]; kk < r.cptr[i + 1]; ++kk) {
      uint32_t s = r.cidx[kk];
      uint32_t e = s >> 1;
      int nb = ((s & 1u) != 0u) ? r.ecL[e] : r.ecR[e];
      bool dry_side = nb >= 0 && r.h[nb] < pc.dry;
      any_dry = any_dry || dry_side;
      float kf = dry_side ? 2.0f : 1.0f;
      if (F) lam = std::fma(std::fma(kf, c, u), r.elen[e], lam);
      else lam += (u + kf * c) * r.elen[e];
      perim += r.elen[e];
    }
    if (F && all_wet[i >> 8] && !any_dry) lam = (u + c) * perim;
    best = std::min(best, cell_bound(r, pc, i, lam));
  }
  r.dtbuf[1] = std::min(r.dtbuf[1], best);
}

void block_mark(cref& r, bool all_active) {
  for (uint32_t b = 0; b < r.NB; ++b) {
    if (all_active) {
      r.blk[3 * b + 0] = 0u;
      r.blk[3 * b + 1] = 1u;
      continue;
    }
    uint32_t first = b << 8, last = std::min(first + kBlk, r.N), live = 0u;
    for (uint32_t i = first; i < last; ++i)
      if (r.h[i] >= r.pc.dry || r.hu[i] != 0.0f || r.hv[i] != 0.0f) { live = 1u; break; }
    r.blk[3 * b + 0] = live;
  }
}

void dilate(cref& r, bool force) {
  if (!force && r.dtbuf[0] == 0u) return;
  std::vector<uint32_t> act(r.NB);
  for (uint32_t b = 0; b < r.NB; ++b) {
    uint32_t v = r.blk[3 * b + 0] | r.blk[3 * b + 2];
    for (uint32_t k = r.bptr[b]; v == 0u && k < r.bptr[b + 1]; ++k)
      v = r.blk[3 * r.bidx[k] + 0] | r.blk[3 * r.bidx[k] + 2];
    act[b] = v;
  }
  for (uint32_t b = 0; b < r.NB; ++b) r.blk[3 * b + 1] = act[b];
}

void cfl_prologue(cref& r) {
  const Pc pc = r.pc;
  const fastm::K kk = fastm::consts(pc);
#pragma omp parallel for schedule(static)
  for (int64_t ee = 0; ee < (int64_t)r.E; ++ee) {
    uint32_t eid = (uint32_t)ee;
    int cL = r.ecL[eid], cR = r.ecR[eid];
    float nx_ = r.enx[eid], ny_ = r.eny[eid];
    float hL = r.h[cL], huL = r.hu[cL], hvL = r.hv[cL];
    float hR, huR, hvR;
    if (cR == -1) {
      hR = hL;
      huR = huL - 2.0f * (huL * nx_ + hvL * ny_) * nx_;
      hvR = hvL - 2.0f * (huL * nx_ + hvL * ny_) * ny_;
    } else if (cR < -1) {
      const uint32_t k = (uint32_t)(-2 - cR);
      ghost_state(pc, r.okind[k], r.oval[k], hL, huL, hvL, nx_, ny_, hR, huR, hvR);
    } else {
      hR = r.h[cR]; huR = r.hu[cR]; hvR = r.hv[cR];
    }
    float contrib = 0.0f;
    if (hL >= pc.dry || hR >= pc.dry) {
      if (r.fast) {
        const float loL = load_lo(r, 0, cL);
        const fastm::Cell L = fastm::cell(kk, hL, huL, hvL, loL);
        const fastm::Cell R = (cR == -1) ? L : fastm::cell(kk, hR, huR, hvR, cR >= 0 ? load_lo(r, 0, cR) : loL);
        const float dzL = r.edz[2 * eid], dzR = r.edz[2 * eid + 1];
        contrib = fastm::wave_bound(kk, L, R, cR == -1, (dzL > 0.0f) ? dzL : -dzR, nx_, ny_) * r.elen[eid];
      } else {
        contrib = wave_bound(pc, hL, huL, hvL, r.edz[2 * eid], hR, huR, hvR, r.edz[2 * eid + 1], nx_, ny_) *
                  r.elen[eid];
      }
    }
    r.ef[(size_t)kWords * eid + kCfl] = contrib;
  }
  uint32_t best = 0xffffffffu;
#pragma omp parallel for schedule(static) reduction(min : best)
  for (int64_t ii = 0; ii < (int64_t)r.N; ++ii) {
    uint32_t i = (uint32_t)ii;
    float lam = 0.0f;
    for (uint32_t k = r.cptr[i]; k < r.cptr[i + 1]; ++k) lam += r.ef[(size_t)kWords * (r.cidx[k] >> 1) + kCfl];
    best = std::min(best, cell_bound(r, pc, i, lam));
  }
  r.dtbuf[0] = std::min(r.dtbuf[0], best);
}

void step_control(cref& r, int mode) {
  Ctrl& c = r.ctrl;
  float raw;
  if (mode == 0) {
    raw = bitsf(r.dtbuf[0]);
    c.last_raw = raw;
  } else if (mode == 1) {
    uint32_t done = c.dt_ticks;
    c.now_ticks += (uint64_t)done;
    if (done > 0u) {
      c.steps += 1u;
      raw = bitsf(r.dtbuf[1]);
      c.last_raw = raw;
    } else {
      raw = c.last_raw;
    }
  } else {
    raw = c.last_raw;
  }
  uint64_t remaining = (c.stop_ticks > c.now_ticks) ? c.stop_ticks - c.now_ticks : 0u;
  uint32_t ticks;
  if (remaining == 0u || raw < 1e-10f) ticks = 0u;
  else if (raw >= 1e10f) ticks = c.dry_dt_ticks;
  else {
    float cap = (float)c.dt_max_ticks / kTicks;
    ticks = (uint32_t)(gmin(raw * c.cfl_safety, cap) * kTicks);
  }
  ticks = std::min(ticks, c.dt_max_ticks);
  if ((uint64_t)ticks > remaining) ticks = (uint32_t)remaining;
  c.dt_ticks = ticks;
  r.dtbuf[0] = fbits((float)ticks / kTicks);
  r.dtbuf[1] = fbits(kSentinel);
}

}  // namespace

extern "C" {

cref* cref_create(uint32_t N, uint32_t E, const int32_t* ecL, const int32_t* ecR, const float* nx,
                  const float* ny, const float* len, const float* dz2, const float* area,
                  const uint32_t* cptr, const uint32_t* cidx, const uint32_t* erange,
                  const uint32_t* bptr, const uint32_t* bidx) {
  cref* r = new cref();
  r->N = N;
  r->E = E;
  r->NB = (N + kBlk - 1) / kBlk;
  r->NWG = (E + kWg - 1) / kWg;
  r->ecL.assign(ecL, ecL + E);
  r->ecR.assign(ecR, ecR + E);
  r->enx.assign(nx, nx + E);
  r->eny.assign(ny, ny + E);
  r->elen.assign(len, len + E);
  r->edz.assign(dz2, dz2 + 2 * (size_t)E);
  r->area.assign(area, area + N);
  r->ia.resize(N);
  for (uint32_t i = 0; i < N; ++i) r->ia[i] = 1.0f / area[i];
  r->cptr.assign(cptr, cptr + N + 1);
  r->cidx.assign(cidx, cidx + cptr[N]);
  r->erange.assign(erange, erange + 2 * (size_t)r->NWG);
  r->bptr.assign(bptr, bptr + r->NB + 1);
  r->bidx.assign(bidx, bidx + bptr[r->NB]);
  for (auto* v : {&r->h, &r->hu, &r->hv, &r->h1, &r->hu1, &r->hv1, &r->hlo, &r->h1lo, &r->clamp}) v->assign(N, 0.0f);
  r->nman.assign(N, 0.03f);
  r->ef.assign((size_t)kWords * E, 0.0f);
  r->touched.assign(N, 0);
  r->blk.assign(3 * (size_t)r->NB, 0u);
  // Open edges (ecR = -2 - k) start closed, in one stretch.
  for (uint32_t e = 0; e < E; ++e) {
    if (ecR[e] >= -1) continue;
    const uint32_t k = (uint32_t)(-2 - ecR[e]);
    if (k >= r->oedge.size()) r->oedge.resize(k + 1, 0u);
    r->oedge[k] = e;
  }
  const uint32_t n_open = (uint32_t)r->oedge.size();
  r->okind.assign(n_open, kOpenWall);
  r->oval.assign(n_open, 0.0f);
  r->sptr = {0u, n_open};
  r->tally.assign(n_open ? 4 : 0, 0.0f);
  return r;
}

void cref_destroy(cref* r) { delete r; }

void cref_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int cref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void cref_set_params(cref* r, float g, float dry, float cfl, uint32_t flags) {
  r->pc = Pc{g, dry, cfl};
  r->compensated = (flags & 2u) != 0;
  r->active_set = (flags & 4u) != 0;
  r->track_clamp = (flags & 8u) != 0;
  r->fast = (flags & 16u) == 0;
}

void cref_set_state(cref* r, const float* h, const float* hu, const float* hv, const float* hlo) {
  r->h.assign(h, h + r->N);
  r->hu.assign(hu, hu + r->N);
  r->hv.assign(hv, hv + r->N);
  if (hlo) r->hlo.assign(hlo, hlo + r->N);
  else r->hlo.assign(r->N, 0.0f);
  r->h1lo.assign(r->N, 0.0f);
  std::fill(r->tally.begin(), r->tally.end(), 0.0f);
}

// Stretch s holds open edges [ptr[s], ptr[s + 1]); clears the tally.
void cref_set_open_stretches(cref* r, uint32_t n, const uint32_t* ptr) {
  r->sptr.assign(ptr, ptr + n + 1);
  r->tally.assign(4 * (size_t)n, 0.0f);
}

// Condition and value per open edge; a block whose open edge can bring water
// in (inflow or stage above zero) is forced live (gpu_resident_batch.py:447-458).
void cref_set_open_conditions(cref* r, const uint32_t* kinds, const float* values) {
  const size_t n = r->oedge.size();
  r->okind.assign(kinds, kinds + n);
  r->oval.assign(values, values + n);
  for (uint32_t b = 0; b < r->NB; ++b) r->blk[3 * b + 2] = 0u;
  for (size_t k = 0; k < n; ++k)
    if ((kinds[k] == kOpenInflow || kinds[k] == kOpenStage) && values[k] > 0.0f)
      r->blk[3 * ((uint32_t)r->ecL[r->oedge[k]] >> 8) + 2] = 1u;
}

uint32_t cref_open_count(const cref* r) { return (uint32_t)r->oedge.size(); }

// Out and in volume per stretch as (hi, lo) float pairs: 4 words per stretch.
void cref_open_tally(const cref* r, float* out4) { std::copy(r->tally.begin(), r->tally.end(), out4); }

void cref_set_manning(cref* r, const float* n) { r->nman.assign(n, n + r->N); }

void cref_set_sources(cref* r, uint32_t n, const uint32_t* cell, const float* rate) {
  r->src_cell.assign(cell, cell + n);
  r->src_rate.assign(rate, rate + n);
}

void cref_ctrl_write(cref* r, const void* ctrl48) { memcpy(&r->ctrl, ctrl48, 48); }
void cref_ctrl_read(const cref* r, void* ctrl48) { memcpy(ctrl48, &r->ctrl, 48); }
void cref_dt_words(const cref* r, uint32_t* out4) { memcpy(out4, r->dtbuf, 16); }

// gpu_resident_batch.py:567-591
void cref_prologue(cref* r) {
  block_mark(*r, !r->active_set);
  if (r->active_set) dilate(*r, true);
  r->dtbuf[0] = fbits(kSentinel);
  cfl_prologue(*r);
  step_control(*r, 0);
}

void cref_retarget(cref* r) { step_control(*r, 2); }

// step_control alone on a hand-set control block and dt words (dt4[0] the
// prologue bound, dt4[1] the next-step bound), both updated in place; the
// same probe as the backend's inu_probe_clock.
void cref_probe_clock(int mode, void* ctrl48, uint32_t* dt4) {
  cref r;
  memcpy(&r.ctrl, ctrl48, 48);
  memcpy(r.dtbuf, dt4, 16);
  step_control(r, mode);
  memcpy(ctrl48, &r.ctrl, 48);
  memcpy(dt4, r.dtbuf, 16);
}

// gpu_resident_batch.py:489-513
void cref_steps(cref* r, uint32_t n) {
  for (uint32_t k = 0; k < n; ++k) {
    for (int stage = 0; stage < 2; ++stage) {
      flux(*r, stage);
      open_tally(*r);
      update(*r, stage);
    }
    if (!r->src_cell.empty()) source(*r);
    if (r->active_set) dilate(*r, false);
    step_control(*r, 1);
  }
}

// Fields as the backend numbers them: h hu hv h_lo h1 hu1 hv1 h1_lo clamp,
// then per-block wet, active, forced.
void cref_download(const cref* r, int which, void* out) {
  const std::vector<float>* f[] = {&r->h, &r->hu, &r->hv, &r->hlo, &r->h1, &r->hu1, &r->hv1, &r->h1lo, &r->clamp};
  if (which >= 0 && which <= 8) {
    memcpy(out, f[which]->data(), 4 * (size_t)r->N);
  } else if (which >= 9 && which <= 11) {
    uint32_t* o = static_cast<uint32_t*>(out);
    for (uint32_t b = 0; b < r->NB; ++b) o[b] = r->blk[3 * b + (which - 9)];
  }
}

// Overwrites one cell field (the backend's inu_upload): h hu hv h_lo h1 hu1
// hv1 h1_lo clamp.
void cref_upload(cref* r, int which, const void* in) {
  std::vector<float>* f[] = {&r->h, &r->hu, &r->hv, &r->hlo, &r->h1, &r->hu1, &r->hv1, &r->h1lo, &r->clamp};
  if (which >= 0 && which <= 8) memcpy(f[which]->data(), in, 4 * (size_t)r->N);
}

// Same probe as the backend's inu_probe_open: case k = (h, hu, hv, nx, ny, fh)
// -> ghost (h, hu, hv) and boundary mass flux.
void cref_probe_open(uint32_t n, const float* cases6, const uint32_t* kinds, const float* values, float* out4,
                     float g, float dry) {
  Pc pc{g, dry, 0.8f};
  for (uint32_t k = 0; k < n; ++k) {
    const float* c = cases6 + 6 * (size_t)k;
    float* o = out4 + 4 * (size_t)k;
    ghost_state(pc, kinds[k], values[k], c[0], c[1], c[2], c[3], c[4], o[0], o[1], o[2]);
    o[3] = boundary_mass_flux(kinds[k], values[k], c[5]);
  }
}

// Same probe as the backend's inu_probe.
void cref_probe(uint32_t n, const float* pairs, float* out7, float g, float dry) {
  Pc pc{g, dry, 0.8f};
  const fastm::K k = fastm::consts(pc);
  for (uint32_t j = 0; j < n; ++j) {
    float x = pairs[2 * j], y = pairs[2 * j + 1];
    float s, e;
    two_sum(x, y, s, e);
    float* o = out7 + 7 * (size_t)j;
    o[0] = e;
    o[1] = x / y;
    o[2] = std::sqrt(std::fabs(x));
    o[3] = (x > 0.0f) ? pow43(x) : 0.0f;
    o[4] = (x > 0.0f) ? film_drag(pc, x, y) : 0.0f;
    o[5] = (x > 0.0f) ? fastm::rcbrt(x) : 0.0f;
    o[6] = (x >= dry && y >= 0.0f) ? fastm::film(k, x, y) : 0.0f;  // its domain
  }
}

}  // extern "C"
