// Exact FP64 predicates shared by every refinement kernel (and by host code that must agree with them bit for bit).
// The engine is compiled with -ffp-contract=off: every product and sum below is rounded exactly as written, and the
// only fused operations are the explicit fma calls of orient().
//
// Coordinates reaching these predicates are origin-relative (v - origin of the pair's reference geometry). For
// geometries a few degrees apart the subtractions that produce them are exact (Sterbenz), so a vertex shared by two
// geometries has bit-identical coordinates in every pair frame and the determinant below has an exact sign.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

namespace gsj {

#define GSJ_HD __host__ __device__ __forceinline__

// (b - a) x (c - a) with the rounding error of the c*d product recovered by fma (Kahan's 2x2 determinant).
// reference: engine/engine.cuh:283-289
GSJ_HD double orient(double ax, double ay, double bx, double by, double cx, double cy) {
  const double a = bx - ax, b = cy - ay, c = by - ay, d = cx - ax;
  const double cd = c * d;
  const double cd_err = fma(-c, d, cd);
  const double ab_m_cd = fma(a, b, -cd);
  return ab_m_cd + cd_err;
}

// closed bbox membership of a collinear point (engine/engine.cuh:290-292)
GSJ_HD bool onseg(double ax, double ay, double bx, double by, double px, double py) {
  return px >= fmin(ax, bx) && px <= fmax(ax, bx) && py >= fmin(ay, by) && py <= fmax(ay, by);
}

// Is p left of a->b? Exactly-collinear p is displaced by (eps, eps^2) (simulation of simplicity), which makes the
// answer consistent between an edge and its reverse. reference: engine/engine.cuh:298-304
GSJ_HD bool left_of(double ax, double ay, double bx, double by, double px, double py) {
  const double o = orient(ax, ay, bx, by, px, py);
  if (o != 0) return o > 0;
  const double dy = by - ay;
  if (dy != 0) return dy < 0;
  return bx > ax;
}

// closed segments p0p1 and q0q1 share a point (touching counts). reference: engine/engine.cuh:307-316
GSJ_HD bool seg_isect(double p0x, double p0y, double p1x, double p1y, double q0x, double q0y, double q1x, double q1y) {
  const double d1 = orient(q0x, q0y, q1x, q1y, p0x, p0y), d2 = orient(q0x, q0y, q1x, q1y, p1x, p1y);
  const double d3 = orient(p0x, p0y, p1x, p1y, q0x, q0y), d4 = orient(p0x, p0y, p1x, p1y, q1x, q1y);
  const bool cross_q = (d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0);
  const bool cross_p = (d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0);
  if (cross_q && cross_p) return true;
  if (d1 == 0 && onseg(q0x, q0y, q1x, q1y, p0x, p0y)) return true;
  if (d2 == 0 && onseg(q0x, q0y, q1x, q1y, p1x, p1y)) return true;
  if (d3 == 0 && onseg(p0x, p0y, p1x, p1y, q0x, q0y)) return true;
  if (d4 == 0 && onseg(p0x, p0y, p1x, p1y, q1x, q1y)) return true;
  return false;
}

// evidence bits and DE-9IM encoding (engine/engine.cuh:334-338, 845-856)
constexpr uint32_t REL_COIN_SAME = 1;   // real boundaries coincide, interiors on the same side
constexpr uint32_t REL_COIN_OPP = 2;    // real boundaries coincide, interiors on opposite sides
constexpr uint32_t REL_TOUCH = 4;       // real boundaries meet
constexpr int MAXS = 24;                // split points kept per edge
constexpr int MAXO = 8;                 // collinear overlaps kept per edge

// DE-9IM of two areal geometries from boundary evidence e = (A inside B, A on B, B inside A, B on A) and the real
// boundary lengths LA, LB; cells II IB IE BI BB BE EI EB EE at 2 bits each (0 = F, 1 = '0', 2 = '1', 3 = '2')
GSJ_HD uint32_t de9im_code(double ai, double ao, double bi, double bo, uint32_t bits, double LA, double LB) {
  const double ta = 1e-9 * LA, tb = 1e-9 * LB;
  const bool aIn = ai > ta, aOut = (LA - ai - ao) > ta, bIn = bi > tb, bOut = (LB - bi - bo) > tb;
  const bool same = bits & REL_COIN_SAME, opp = bits & REL_COIN_OPP, touch = bits & REL_TOUCH;
  const uint32_t II = (aIn || bIn || same) ? 3 : 0, IB = bIn ? 2 : 0, IE = (aOut || bIn || opp) ? 3 : 0;
  const uint32_t BI = aIn ? 2 : 0, BB = (same || opp) ? 2 : (touch ? 1 : 0), BE = aOut ? 2 : 0;
  const uint32_t EI = (bOut || aIn || opp) ? 3 : 0, EB = bOut ? 2 : 0, EE = 3;
  return II | IB << 2 | IE << 4 | BI << 6 | BB << 8 | BE << 10 | EI << 12 | EB << 14 | EE << 16;
}

}  // namespace gsj
