// nm_fp.h -- IEEE arithmetic in a selectable rounding mode, computed in
// the default round-to-nearest-even hardware mode.
//
// The reference executes its host path with FE_DOWNWARD left set by
// SensorData::updatePose (include/nanomap/sensor/SensorData.h:42), so the
// hit/miss traversal and the FP32 log-odds sums round toward -inf.  On
// gfx950 the MODE register can select that rounding, but the compiler's
// mode-register pass resets the FP64 rounding bits around its own div/sqrt
// sequences without restoring them.  So outside the two per-step DDA loops
// round-down is emulated exactly here (those loops hold no divide or square
// root; nm_cast.hip sets the wave's FP64 mode to toward -inf around them
// and back to nearest after, and their crossing-time adds are then the
// hardware add):
//   add: TwoSum error term; result stepped down one ulp if the error < 0;
//   mul: FMA residual a*b - p; stepped down if < 0;
//   div: FMA residual a - q*b of the correctly rounded quotient;
//   sqrt: FMA residual x - s*s of the correctly rounded root.
// Each residual is exact, so each result equals the one IEEE operation
// rounded toward -inf.  RN instantiations are the plain operations.
//
// Build with -ffp-contract=off: a contracted a*b+c changes the result.
#ifndef NM_FP_H
#define NM_FP_H

#include <limits>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NM_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define NM_HD inline
#endif

namespace nm {

enum Rounding : int { kNearest = 0, kDown = 1 };

NM_HD uint64_t bits_of(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
NM_HD double double_of(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}
NM_HD uint32_t bits_of(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
NM_HD float float_of(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

template <class T>
NM_HD T abs_t(T x) {
  return x < T(0) ? -x : x;
}

// Next representable value toward -inf, for finite nonzero x.
NM_HD double next_down(double x) {
  const uint64_t u = bits_of(x);
  return double_of((u >> 63) ? u + 1 : u - 1);
}
NM_HD float next_down(float x) {
  const uint32_t u = bits_of(x);
  return float_of((u >> 31) ? u + 1 : u - 1);
}

NM_HD double fma_(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}
NM_HD float fma_(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmaf(a, b, c);
#else
  return std::fmaf(a, b, c);
#endif
}
NM_HD double sqrt_(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrt(x);
#else
  return std::sqrt(x);
#endif
}
NM_HD float sqrt_(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}

// Largest finite value of T.
template <class T>
NM_HD T max_finite();
template <>
NM_HD float max_finite<float>() {
  return 3.40282347e38f;
}
template <>
NM_HD double max_finite<double>() {
  return 1.7976931348623157e308;
}

// Smallest-magnitude negative value of T (the round-down of a negative
// result below it in magnitude).
template <class T>
NM_HD T neg_tiny() {
  return -std::numeric_limits<T>::denorm_min();
}

// Sum rounded toward -inf.  An exact zero sum of nonzero operands is -0
// in that mode; +0 only when both operands are +0.  A finite sum above the
// largest finite value rounds down to it.
template <class T>
NM_HD T add_down(T a, T b) {
  T s = a + b;
  if (s == T(0)) {
    const bool both_pos_zero = (a == T(0) && b == T(0) &&
                                !(bits_of(a) >> (sizeof(T) * 8 - 1)) &&
                                !(bits_of(b) >> (sizeof(T) * 8 - 1)));
    return both_pos_zero ? T(0) : -T(0);
  }
  if (s > max_finite<T>() && a <= max_finite<T>() && b <= max_finite<T>())
    return max_finite<T>();  // +inf from finite operands
  const T bb = s - a;
  const T err = (a - (s - bb)) + (b - bb);
  return err < T(0) ? next_down(s) : s;
}

// add_down for operands whose sum is known to be finite (no overflow case).
template <class T>
NM_HD T add_down_finite(T a, T b) {
  T s = a + b;
  if (s == T(0)) {
    const bool both_pos_zero = (a == T(0) && b == T(0) &&
                                !(bits_of(a) >> (sizeof(T) * 8 - 1)) &&
                                !(bits_of(b) >> (sizeof(T) * 8 - 1)));
    return both_pos_zero ? T(0) : -T(0);
  }
  const T bb = s - a;
  const T err = (a - (s - bb)) + (b - bb);
  return err < T(0) ? next_down(s) : s;
}

// Round-down sum of a >= 0 and b > 0 (the DDA's next-crossing update),
// branch-free: Fast2Sum on (max, min), then one ulp down if the error is
// negative.  The sum is positive, so the step down is an integer decrement.
NM_HD double add_down_pos(double a, double b) {
  const double s = a + b;
  const double big = a > b ? a : b;
  const double small = a > b ? b : a;
  const double err = small - (s - big);
  return err < 0.0 ? double_of(bits_of(s) - 1) : s;
}

// (An overflowing product or quotient comes out right through next_down:
// the residual against infinity has the sign that steps +inf down to the
// largest finite value and keeps -inf.)
template <class T>
NM_HD T mul_down(T a, T b) {
  const T p = a * b;
  if (p == T(0)) {
    // a nonzero negative product below the smallest subnormal
    const bool neg = (a < T(0)) != (b < T(0));
    return (neg && a != T(0) && b != T(0)) ? neg_tiny<T>() : p;
  }
  const T err = fma_(a, b, -p);
  return err < T(0) ? next_down(p) : p;
}

template <class T>
NM_HD T div_down(T a, T b) {
  const T q = a / b;
  if (b == T(0)) return q;
  if (q == T(0)) {
    const bool neg = (a < T(0)) != (b < T(0));
    return (neg && a != T(0) && !(abs_t(b) > max_finite<T>())) ? neg_tiny<T>() : q;
  }
  const T r = fma_(-q, b, a);  // exact: a - q*b
  const bool q_above = (b > T(0)) ? (r < T(0)) : (r > T(0));
  return q_above ? next_down(q) : q;
}

template <class T>
NM_HD T sqrt_down(T x) {
  const T s = sqrt_(x);
  if (s == T(0)) return s;
  const T r = fma_(-s, s, x);  // exact: x - s*s
  return r < T(0) ? next_down(s) : s;
}

// Mode-dispatched operations; M is a compile-time constant in kernels.
// kFinite: the caller guarantees a finite sum (bounded operands), which
// drops the overflow case from the round-down sum.
template <int M, bool kFinite = false, class T>
NM_HD T add(T a, T b) {
  if (M != kDown) return a + b;
  return kFinite ? add_down_finite(a, b) : add_down(a, b);
}
template <int M, class T>
NM_HD T sub(T a, T b) {
  return M == kDown ? add_down(a, -b) : a - b;
}
template <int M, class T>
NM_HD T mul(T a, T b) {
  return M == kDown ? mul_down(a, b) : a * b;
}
template <int M, class T>
NM_HD T div(T a, T b) {
  return M == kDown ? div_down(a, b) : a / b;
}
template <int M, class T>
NM_HD T sqrt(T x) {
  return M == kDown ? sqrt_down(x) : sqrt_(x);
}

// ---------------------------------------------------------------------
// fold<M>(v, a, k): v_{i+1} = round_M(v_i + a) applied k times (FP32).
//
// Inside one binade [2^e, 2^(e+1)) every float is a multiple of
// u = 2^(e-23), so while the exact sums v_i + a stay in that binade the
// rounded step is the constant d = round_M(a / u) * u.  The loop jumps
// over such runs in one exact multiply-add and takes single rounded steps
// near binade boundaries, near zero and at round-to-nearest ties (where
// the step depends on the parity of v).  Cost: O(number of binades).
// ---------------------------------------------------------------------
NM_HD double floor_(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_floor(x);
#else
  return std::floor(x);
#endif
}
NM_HD double rint_(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rint(x);
#else
  return std::rint(x);
#endif
}
NM_HD double abs_(double x) { return x < 0 ? -x : x; }
// An approximation of 1/x good to a few ulps (callers correct the integer
// quotients they derive from it).
NM_HD double approx_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcp(x);
#else
  return 1.0 / x;
#endif
}

//
// kFinite: v, a and every partial sum are known to be finite (bounded
// inputs, as in the cast kernels); otherwise an infinite or NaN v returns
// at once, since it stays what it is under any further step.
template <int M, bool kFinite = false>
NM_HD float fold(float v, float a, uint32_t k) {
  if (a == 0.0f) return k ? add<M, kFinite>(v, a) : v;
  if (!kFinite && !(abs_((double)v) <= 3.4028234663852886e38))
    return k ? add<M>(v, a) : v;
  // short runs: plain rounded steps are cheaper than the run analysis
  while (k > 0 && k <= 24) {
    v = add<M, kFinite>(v, a);
    --k;
  }
  while (k > 0) {
    if (!kFinite && !(abs_((double)v) <= 3.4028234663852886e38))
      return add<M>(v, a);
    const double av = abs_((double)v), aa = abs_((double)a);
    // Single steps while |v| is small relative to the increment, so every
    // jump below stays clear of zero crossings and subnormals.
    if (av < 2.0 * aa || av < 1e-30) {
      v = add<M, kFinite>(v, a);
      --k;
      continue;
    }
    const uint32_t ex = (bits_of(v) >> 23) & 0xff;  // biased exponent >= 1
    const double lo = double_of((uint64_t)(ex + 896) << 52);      // 2^e
    const double ulp = double_of((uint64_t)(ex + 896 - 23) << 52);  // 2^(e-23)
    const double inv_ulp = double_of((uint64_t)(1173 - ex) << 52);  // 1 / ulp
    const double hi = 2.0 * lo;
    const double q = (double)a * inv_ulp;  // exact: power-of-two scaling
    double dq;
    if (M == kDown) {
      dq = floor_(q);
    } else {
      if (q - floor_(q) == 0.5) {  // tie: step depends on parity of v
        v = add<M, kFinite>(v, a);
        --k;
        continue;
      }
      dq = rint_(q);
    }
    const double d = dq * ulp;  // exact
    const bool away = ((double)v > 0) == (a > 0.0f);
    // Bounds in whole ulps, all exact integers in double: m = |d|/u,
    // qa = |a|/u (exact scaling), and |v| - 2^e, 2^(e+1) - |v| (multiples
    // of u below 2^24 u).  Step i (from |v| -+ i|d|) keeps the exact sum in
    // the binade iff
    //   away:      i m + qa <  (2^(e+1) - |v|)/u  <=>  i m <= that - floor(qa) - 1
    //   toward 0:  i m + qa <= (|v| - 2^e)/u      <=>  i m <= that - ceil(qa)
    // (i m is an integer).  Working in ulps keeps |a| exactly: a difference
    // such as |v| - |a| - 2^e in double drops an |a| below the double ulp of
    // |v| and would let the last step land on 2^e while the exact sum is
    // just below it.
    const double m = abs_(dq);
    const double qa = abs_(q);
    const double lim = away ? (hi - av) * inv_ulp - floor_(qa) - 1.0
                            : (av - lo) * inv_ulp + floor_(-qa);
    int64_t n = 0;
    if (lim >= 0.0) {
      // the first step stays in the binade: with d == 0 it leaves v as it
      // is, and so does every further step
      if (m == 0.0) return v;
      // floor(lim / m) from an approximate reciprocal, then made exact
      int64_t i = (int64_t)floor_(lim * approx_rcp(m));
      if (i < 0) i = 0;
      while (i > 0 && (double)i * m > lim) --i;
      while ((double)(i + 1) * m <= lim) ++i;
      n = i + 1;
    }
    if (n <= 0) {
      // one rounded step across the binade edge; a step that leaves v as
      // it is repeats forever
      const float nv = add<M, kFinite>(v, a);
      --k;
      if (bits_of(nv) == bits_of(v)) return v;
      v = nv;
      continue;
    }
    if ((uint64_t)n > k) n = (int64_t)k;
    v = (float)((double)v + (double)n * d);  // exact: stays on the grid
    k -= (uint32_t)n;
  }
  return v;
}

}  // namespace nm

#endif  // NM_FP_H
