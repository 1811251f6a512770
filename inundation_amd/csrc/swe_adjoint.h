// Reverse-mode derivatives of the shallow-water step, by hand, for the adjoint.
//
// They differentiate the scheme in real arithmetic (the compensated depth pair
// is one depth), as the forward evaluates it in any arithmetic order:
//   - desingularised velocity (reference flux.py:63-66),
//   - Audusse face depth max(0, h - dz) (flux.py:146-169),
//   - Roe speeds blended with Toro's dry-bed speeds (flux.py:75-95),
//   - HLLC with the 1e-12 / 1e-14 guards (flux.py:98-140),
//   - the hydrostatic correction and walls passing no mass (resident.py:148-151, 268-271),
//   - positivity clamp, dry momentum zeroing, film drag and semi-implicit
//     Manning (resident.py:459-483, flux.py:43-54).
// Every branch (wet/dry, clamps, the HLLC case, min/max, guards) follows the
// forward values. Held (no derivative through them):
//   - the weights of the wet/dry speed blend: they are the wet/dry decision
//     spread over one dry_tol, and their slope 1 / dry_tol, met at every
//     front cell crossing the band next to deep water, compounds over a run
//     (on a 250 k-cell inflow flood the exact tangent reaches 1e20 after 2000
//     steps while central differences stay O(10));
//   - the star state of a degenerate star region (|S - SM| at the 1e-14
//     guard, whose stand-in divisor would give slopes of 1e14);
//   - below adj_thin face depth, a side's velocity and celerity (slopes of
//     order 1 / h in a thin layer); its depth still enters.
// Conventions at kinks: the square root of a non-positive argument has
// derivative 0; max(0, x) and clamps pass the derivative at their bounds; a
// bed tie puts the face on the left cell.
#pragma once

#include "swe_math.h"

namespace inu {
namespace adj {

__host__ __device__ __forceinline__ float sq0(float x) { return x > 0.0f ? sqrtf(x) : 0.0f; }

// vel_factor(h) = 2h / (h^2 + max(h^2, e^2)) for h > dry, else 0; and d/dh.
__host__ __device__ __forceinline__ void vf_d(const Phys& p, float h, float& v, float& dv) {
  if (!(h > p.dry)) {
    v = 0.0f;
    dv = 0.0f;
    return;
  }
  const float hh = h * h;
  if (hh >= p.e_film2) {
    v = 1.0f / h;
    dv = -v * v;
  } else {
    const float den = hh + p.e_film2;
    v = 2.0f * h / den;
    dv = 2.0f * (p.e_film2 - hh) / (den * den);
  }
}

// Gradients of the five update words of one edge with respect to its
// inputs. In: the two sides' stage input, the bed steps, the edge geometry
// and lw[0..4] = d loss / d word. Out: gL, gR (d loss / d (h, hu, hv) of each
// side; for a wall the mirrored right side is folded into gL) and gdz (d loss
// / d dzL, dzR). An edge dry on both sides gives zeros.
__host__ __device__ inline void edge_vjp(const Phys& p, bool wall, float hL, float huL, float hvL, float hR,
                                         float huR, float hvR, float dzL, float dzR, float nx, float ny, float len,
                                         const float lw[5], float gL[3], float gR[3], float gdz[2]) {
  for (int k = 0; k < 3; ++k) {
    gL[k] = 0.0f;
    gR[k] = 0.0f;
  }
  gdz[0] = 0.0f;
  gdz[1] = 0.0f;
  if (wall) {
    const float qn = huL * nx + hvL * ny;
    hR = hL;
    huR = huL - 2.0f * qn * nx;
    hvR = hvL - 2.0f * qn * ny;
  }
  if (hL < p.dry && hR < p.dry) return;
  const float g = p.g;
  // ---- forward ----
  float vfL, dvfL, vfR, dvfR;
  vf_d(p, hL, vfL, dvfL);
  vf_d(p, hR, vfR, dvfR);
  const float uL = huL * vfL, vL = hvL * vfL, uR = huR * vfR, vR = hvR * vfR;
  // Face depth max(0, h - dz): the derivative passes at h - dz = 0 (a dry side
  // without a step has face depth h, derivative 1).
  const bool posL = hL - dzL >= 0.0f, posR = hR - dzR >= 0.0f;
  const float hLs = posL ? hL - dzL : 0.0f;
  const float hRs = posR ? hR - dzR : 0.0f;
  const float huLs = hLs * uL, hvLs = hLs * vL, huRs = hRs * uR, hvRs = hRs * vR;
  const float qnL = huLs * nx + hvLs * ny, qtL = -huLs * ny + hvLs * nx;
  const float qnR = huRs * nx + hvRs * ny, qtR = -huRs * ny + hvRs * nx;
  float vfLs, dvfLs, vfRs, dvfRs;
  vf_d(p, hLs, vfLs, dvfLs);
  vf_d(p, hRs, vfRs, dvfRs);
  const float unL = qnL * vfLs, utL = qtL * vfLs, unR = qnR * vfRs, utR = qtR * vfRs;
  // speeds
  const float sqL = sq0(hLs), sqR = sq0(hRs);
  const float cL = p.sqrt_g * sqL, cR = p.sqrt_g * sqR;
  const float den = sqL + sqR;
  const bool roe_ok = den > 1e-12f;
  const float u_roe = roe_ok ? (sqL * unL + sqR * unR) / den : 0.0f;
  const float carg = p.half_g * (hLs + hRs);
  const float c_roe = sq0(carg);
  const float a1 = unL - cL, a2 = u_roe - c_roe;
  const bool pick_a2 = a2 < a1;
  const float sl = pick_a2 ? a2 : a1;
  const float b1 = unR + cR, b2 = u_roe + c_roe;
  const bool pick_b2 = b1 < b2;
  const float sr = pick_b2 ? b2 : b1;
  const float xl = hLs * p.inv_dry - 1.0f, xr = hRs * p.inv_dry - 1.0f;
  const float wl = gmin(gmax(xl, 0.0f), 1.0f), wr = gmin(gmax(xr, 0.0f), 1.0f);
  const float dl = (1.0f - wl) * wr, dr = (1.0f - wr) * wl;
  const float roe = 1.0f - dl - dr;
  const float SL = roe * sl + dl * (unR - 2.0f * cR) + dr * (unL - cL);
  const float SR = roe * sr + dl * (unR + cR) + dr * (unL + 2.0f * cL);
  // contact speed
  const float nA = SL * hRs * (unR - SR) - SR * hLs * (unL - SL);
  const float dA = hRs * (unR - SR) - hLs * (unL - SL);
  const bool sm_ok = fabsf(dA) > 1e-12f;
  const float SM = sm_ok ? nA / dA : 0.0f;
  // A side thinner than adj_thin enters the speeds and fluxes with its
  // velocity and celerity held: u = q / h and sqrt(g h) have slopes of order
  // 1 / h there, which the reverse pass would compound at every front.
  const bool thinL = hLs < p.adj_thin, thinR = hRs < p.adj_thin;
  // ---- reverse ----
  float bhL = 0.0f, bhuL = 0.0f, bhvL = 0.0f, bhR = 0.0f, bhuR = 0.0f, bhvR = 0.0f;
  float bhLs = 0.0f, bhRs = 0.0f, bunL = 0.0f, butL = 0.0f, bunR = 0.0f, butR = 0.0f;
  float bSL = 0.0f, bSR = 0.0f, bSM = 0.0f;
  const float bFh = wall ? 0.0f : lw[0] * len;
  const float bFhu = (lw[1] + lw[3]) * len, bFhv = (lw[2] + lw[4]) * len;
  const float bslg = (lw[1] * nx + lw[2] * ny) * len, bsrg = (lw[3] * nx + lw[4] * ny) * len;
  const float bFhn = bFhu * nx + bFhv * ny, bFht = -bFhu * ny + bFhv * nx;
  // hydrostatic corrections slg = g/2 (hL^2 - hLs^2), srg likewise
  bhL += bslg * g * hL;
  bhLs -= bslg * g * hLs;
  bhR += bsrg * g * hR;
  bhRs -= bsrg * g * hRs;
  // HLLC case
  if (SL >= 0.0f) {
    bhLs += bFh * unL + bFhn * (unL * unL + g * hLs) + bFht * unL * utL;
    bunL += bFh * hLs + bFhn * 2.0f * hLs * unL + bFht * hLs * utL;
    butL += bFht * hLs * unL;
  } else if (!(SM >= 0.0f) && !(SR >= 0.0f)) {
    bhRs += bFh * unR + bFhn * (unR * unR + g * hRs) + bFht * unR * utR;
    bunR += bFh * hRs + bFhn * 2.0f * hRs * unR + bFht * hRs * utR;
    butR += bFht * hRs * unR;
  } else {
    const bool left = SM >= 0.0f;
    const float S = left ? SL : SR, hS = left ? hLs : hRs, unS = left ? unL : unR, utS = left ? utL : utR;
    const float dS0 = S - SM;
    const bool ds_ok = fabsf(dS0) > 1e-14f;
    const float dS = ds_ok ? dS0 : 1e-14f;
    const float k = hS * (S - unS) / dS;
    // F = (hS unS + S (k - hS), hS unS^2 + g/2 hS^2 + S (k SM - hS unS), hS unS utS + S utS (k - hS))
    float bhS = bFh * (unS - S) + bFhn * (unS * unS + g * hS - S * unS) + bFht * (unS * utS - S * utS);
    float bunS = bFh * hS + bFhn * (2.0f * hS * unS - S * hS) + bFht * hS * utS;
    const float butS = bFht * (hS * unS + S * (k - hS));
    float bS = bFh * (k - hS) + bFhn * (k * SM - hS * unS) + bFht * utS * (k - hS);
    const float bk = bFh * S + bFhn * S * SM + bFht * S * utS;
    bSM += bFhn * S * k;
    // k = hS (S - unS) / dS, dS = S - SM; a guarded (degenerate) star is held.
    if (ds_ok) {
      bhS += bk * (S - unS) / dS;
      bunS -= bk * hS / dS;
      bS += bk * (hS - k) / dS;
      bSM += bk * k / dS;
    }
    if (left) {
      bhLs += bhS;
      bunL += bunS;
      butL += butS;
      bSL += bS;
    } else {
      bhRs += bhS;
      bunR += bunS;
      butR += butS;
      bSR += bS;
    }
  }
  // SM = nA / dA
  if (sm_ok) {
    const float bn = bSM / dA, bd = -bSM * SM / dA;
    bSL += bn * (hRs * (unR - SR) + SR * hLs) + bd * hLs;
    bSR += bn * (-SL * hRs - hLs * (unL - SL)) - bd * hRs;
    bhRs += bn * SL * (unR - SR) + bd * (unR - SR);
    bhLs += -bn * SR * (unL - SL) - bd * (unL - SL);
    bunR += bn * SL * hRs + bd * hRs;
    bunL += -bn * SR * hLs - bd * hLs;
  }
  // speeds: SL = roe sl + dl (unR - 2 cR) + dr (unL - cL), SR = roe sr + dl (unR + cR) + dr (unL + 2 cL)
  const float bsl = bSL * roe, bsr = bSR * roe;
  bunR += (bSL + bSR) * dl;
  bunL += (bSL + bSR) * dr;
  float bcR = -2.0f * bSL * dl + bSR * dl;
  float bcL = -bSL * dr + 2.0f * bSR * dr;
  // The blend weights wl, wr are held (see the header comment).
  float bu_roe = 0.0f, bc_roe = 0.0f;
  if (pick_a2) {
    bu_roe += bsl;
    bc_roe -= bsl;
  } else {
    bunL += bsl;
    bcL -= bsl;
  }
  if (pick_b2) {
    bu_roe += bsr;
    bc_roe += bsr;
  } else {
    bunR += bsr;
    bcR += bsr;
  }
  float bsqL = p.sqrt_g * bcL, bsqR = p.sqrt_g * bcR;
  if (roe_ok) {
    bsqL += bu_roe * (unL - u_roe) / den;
    bsqR += bu_roe * (unR - u_roe) / den;
    bunL += bu_roe * sqL / den;
    bunR += bu_roe * sqR / den;
  }
  if (carg > 0.0f) {
    const float t = bc_roe * p.half_g * 0.5f / c_roe;
    if (!thinL) bhLs += t;
    if (!thinR) bhRs += t;
  }
  if (hLs > 0.0f && !thinL) bhLs += bsqL * 0.5f / sqL;
  if (hRs > 0.0f && !thinR) bhRs += bsqR * 0.5f / sqR;
  if (thinL) {
    bunL = 0.0f;
    butL = 0.0f;
  }
  if (thinR) {
    bunR = 0.0f;
    butR = 0.0f;
  }
  // face velocities un = qn vf(hs), ut = qt vf(hs)
  const float bqnL = bunL * vfLs, bqtL = butL * vfLs;
  const float bqnR = bunR * vfRs, bqtR = butR * vfRs;
  bhLs += (bunL * qnL + butL * qtL) * dvfLs;
  bhRs += (bunR * qnR + butR * qtR) * dvfRs;
  // face momenta q = rot(hs * (u, v))
  const float bhuLs = bqnL * nx - bqtL * ny, bhvLs = bqnL * ny + bqtL * nx;
  const float bhuRs = bqnR * nx - bqtR * ny, bhvRs = bqnR * ny + bqtR * nx;
  bhLs += bhuLs * uL + bhvLs * vL;
  bhRs += bhuRs * uR + bhvRs * vR;
  const float buL = bhuLs * hLs, bvL = bhvLs * hLs, buR = bhuRs * hRs, bvR = bhvRs * hRs;
  bhuL += buL * vfL;
  bhvL += bvL * vfL;
  bhL += (buL * huL + bvL * hvL) * dvfL;
  bhuR += buR * vfR;
  bhvR += bvR * vfR;
  bhR += (buR * huR + bvR * hvR) * dvfR;
  // face depths hs = max(0, h - dz)
  float bdzL = 0.0f, bdzR = 0.0f;
  if (posL) {
    bhL += bhLs;
    bdzL = -bhLs;
  }
  if (posR) {
    bhR += bhRs;
    bdzR = -bhRs;
  }
  if (wall) {
    // hR = hL, (huR, hvR) = (huL, hvL) - 2 (huL nx + hvL ny) (nx, ny); walls have no step
    gL[0] = bhL + bhR;
    gL[1] = bhuL + bhuR * (1.0f - 2.0f * nx * nx) - bhvR * 2.0f * nx * ny;
    gL[2] = bhvL - bhuR * 2.0f * nx * ny + bhvR * (1.0f - 2.0f * ny * ny);
    return;
  }
  gL[0] = bhL;
  gL[1] = bhuL;
  gL[2] = bhvL;
  gR[0] = bhR;
  gR[1] = bhuR;
  gR[2] = bhvR;
  gdz[0] = bdzL;
  gdz[1] = bdzR;
}

// Reverse of a cell's stage update after the sums: from d loss / d (h, hu,
// hv) out of the stage to d loss / d (h', hu', hv') after the update (before
// clamp's cut and friction), with the Manning gradient. (hp, hup, hvp) is the
// updated state after the clamp and the dry zeroing, before friction; clamped
// is whether h - dt/A sum_h was negative.
__host__ __device__ inline void friction_vjp(const Phys& p, float hp, float hup, float hvp, bool clamped, float n,
                                             float dt, float bh, float bhu, float bhv, float& rh, float& rhu,
                                             float& rhv, float& gn) {
  gn = 0.0f;
  rh = clamped ? 0.0f : bh;
  rhu = 0.0f;
  rhv = 0.0f;
  if (!(hp >= p.dry)) return;  // dry: momentum zeroed, no friction
  const float g = p.g;
  // film drag of a thin film
  float drag = 1.0f, ddrag = 0.0f;
  if (hp < p.e_film) {
    const float s = p.sqrt_g_e;  // sqrt(g / e)
    const float r = (p.e_film2 - hp * hp) / (2.0f * hp * hp) * s;
    drag = 1.0f / (1.0f + dt * r);
    const float dr = -p.e_film2 / (hp * hp * hp) * s;
    ddrag = -dt * dr * drag * drag;
  }
  const float hu2 = hup * drag, hv2 = hvp * drag;
  const float u = hu2 / hp, v = hv2 / hp;
  const float spd = sq0(u * u + v * v);
  const float h43 = powf(hp, 4.0f / 3.0f);
  const float Sf = g * n * n * spd / h43;
  const float den = 1.0f + dt * Sf;
  const float huo = hu2 / den, hvo = hv2 / den;
  const float bden = -(bhu * huo + bhv * hvo) / den;
  float bhu2 = bhu / den, bhv2 = bhv / den;
  const float bSf = bden * dt;
  if (n > 0.0f) gn = bSf * 2.0f * Sf / n;
  float bhp = bSf * (-4.0f / 3.0f) * Sf / hp;
  if (spd > 0.0f) {
    const float bspd = bSf * Sf / spd;
    const float bu = bspd * u / spd, bv = bspd * v / spd;
    bhu2 += bu / hp;
    bhv2 += bv / hp;
    bhp -= (bu * u + bv * v) / hp;
  }
  rhu = bhu2 * drag;
  rhv = bhv2 * drag;
  bhp += (bhu2 * hup + bhv2 * hvp) * ddrag;
  if (!clamped) rh += bhp;
}

}  // namespace adj
}  // namespace inu
